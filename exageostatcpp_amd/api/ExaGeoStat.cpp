// ExaGeoStat.cpp -- C++ host API of the MI355X likelihood path (see ExaGeoStat.hpp).
#include "ExaGeoStat.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

using exageostat::common::Dimension;

namespace exageostat {

void CheckStatus(int aStatus, const char *aWhere) {
    if (aStatus != EGS_SUCCESS) throw std::runtime_error(std::string(aWhere) + " failed: " + egs_last_error());
}

namespace {

template<typename T>
std::vector<double> ToDouble(const T *a, size_t n) {
    return a ? std::vector<double>(a, a + n) : std::vector<double>();
}

template<typename T>
std::vector<double> ToDouble(const std::vector<T> &a) {
    return std::vector<double>(a.begin(), a.end());
}

double Seconds() {
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

bool Verbose() { return configurations::Configurations::GetVerbosity() != common::QUIET_MODE; }

// Registers the covariance kernels under their reference names
// (UnivariateMaternStationary.cpp:35-36, BivariateMaternParsimonious.cpp:35-36).
template<typename T>
bool RegisterKernels() {
    using plugins::PluginRegistry;
    PluginRegistry<kernels::Kernel<T>>::Add("UnivariateMaternStationary",
                                            []() -> kernels::Kernel<T> * { return new kernels::UnivariateMaternStationary<T>(); });
    PluginRegistry<kernels::Kernel<T>>::Add("BivariateMaternParsimonious",
                                            []() -> kernels::Kernel<T> * { return new kernels::BivariateMaternParsimonious<T>(); });
    auto &params = kernels::KernelsConfigurations::GetParametersNumberKernelMap();
    params["UnivariateMaternStationary"] = 3;
    params["BivariateMaternParsimonious"] = 6;
    return true;
}
const bool kKernelsRegistered = RegisterKernels<double>() && RegisterKernels<float>();

template<typename T>
std::unique_ptr<kernels::Kernel<T>> CreateKernel(configurations::Configurations &aConfigurations) {
    kernels::Kernel<T> *k = plugins::PluginRegistry<kernels::Kernel<T>>::Create(aConfigurations.GetKernelName(),
                                                                                 aConfigurations.GetTimeSlot());
    if (!k) throw std::domain_error("Unknown kernel: " + aConfigurations.GetKernelName());
    return std::unique_ptr<kernels::Kernel<T>>(k);
}

egs_context_t g_context = nullptr;

}  // namespace

// ---------------------------------------------------------------- configurations
namespace configurations {

common::Verbose Configurations::mVerbosity = common::STANDARD_MODE;

Configurations::Configurations() { (void) kKernelsRegistered; }

void Configurations::SetTolerance(double aTolerance) { mTolerance = std::pow(10.0, -aTolerance); }

void Configurations::InitTheta(std::vector<double> &aTheta, int aSize) {
    if (aTheta.empty()) aTheta.assign(aSize, -1.0);
    else if (static_cast<int>(aTheta.size()) < aSize) aTheta.resize(aSize, 0.0);
}

void Configurations::InitializeAllTheta() {
    if (mIsThetaInit) return;
    auto &map = kernels::KernelsConfigurations::GetParametersNumberKernelMap();
    auto it = map.find(mKernelName);
    if (it == map.end()) throw std::domain_error("Unknown kernel: " + mKernelName);
    const int np = it->second;
    InitTheta(mInitialTheta, np);
    InitTheta(mLowerBounds, np);
    InitTheta(mUpperBounds, np);
    mStartingTheta = mLowerBounds;
    InitTheta(mEstimatedTheta, np);
    for (int i = 0; i < np; ++i) {
        if (mEstimatedTheta[i] != -1) {
            mLowerBounds[i] = mEstimatedTheta[i];
            mUpperBounds[i] = mEstimatedTheta[i];
            mStartingTheta[i] = mEstimatedTheta[i];
        }
    }
    mIsThetaInit = true;
}

}  // namespace configurations

// ---------------------------------------------------------------- kernels
namespace kernels {

std::unordered_map<std::string, int> &KernelsConfigurations::GetParametersNumberKernelMap() {
    static std::unordered_map<std::string, int> mKernelParametersNumbers;
    return mKernelParametersNumbers;
}

namespace {
template<typename T>
void DeviceTile(int aKernelId, int aParams, T *apMatrixA, int aRows, int aCols, int aRowOffset, int aColOffset,
                dataunits::Locations<T> &aLocation1, dataunits::Locations<T> &aLocation2, const T *apTheta,
                int aMetric) {
    const size_t n1 = aLocation1.GetSize(), n2 = aLocation2.GetSize();
    auto x1 = ToDouble(aLocation1.GetLocationX(), n1), y1 = ToDouble(aLocation1.GetLocationY(), n1);
    auto z1 = ToDouble(aLocation1.GetLocationZ(), n1);
    auto x2 = ToDouble(aLocation2.GetLocationX(), n2), y2 = ToDouble(aLocation2.GetLocationY(), n2);
    auto z2 = ToDouble(aLocation2.GetLocationZ(), n2);
    auto th = ToDouble(apTheta, aParams);
    std::vector<double> A(static_cast<size_t>(aRows) * aCols);
    CheckStatus(egs_covariance_tile(ExaGeoStatHardware::GetContext(), aKernelId, A.data(), aRows, aCols, aRowOffset,
                                    aColOffset, x1.data(), y1.data(), z1.empty() ? nullptr : z1.data(), (int64_t) n1,
                                    x2.data(), y2.data(), z2.empty() ? nullptr : z2.data(), (int64_t) n2, th.data(),
                                    aMetric),
                "GenerateCovarianceMatrix");
    for (size_t i = 0; i < A.size(); ++i) apMatrixA[i] = static_cast<T>(A[i]);
}
}  // namespace

template<typename T>
void UnivariateMaternStationary<T>::GenerateCovarianceMatrix(T *apMatrixA, const int &aRowsNumber,
                                                             const int &aColumnsNumber, const int &aRowOffset,
                                                             const int &aColumnOffset,
                                                             dataunits::Locations<T> &aLocation1,
                                                             dataunits::Locations<T> &aLocation2,
                                                             dataunits::Locations<T> &, T *apLocalTheta,
                                                             const int &aDistanceMetric) {
    DeviceTile<T>(GetDeviceKernelId(), this->mParametersNumber, apMatrixA, aRowsNumber, aColumnsNumber, aRowOffset,
                  aColumnOffset, aLocation1, aLocation2, apLocalTheta, aDistanceMetric);
}

template<typename T>
void BivariateMaternParsimonious<T>::GenerateCovarianceMatrix(T *apMatrixA, const int &aRowsNumber,
                                                              const int &aColumnsNumber, const int &aRowOffset,
                                                              const int &aColumnOffset,
                                                              dataunits::Locations<T> &aLocation1,
                                                              dataunits::Locations<T> &aLocation2,
                                                              dataunits::Locations<T> &, T *apLocalTheta,
                                                              const int &aDistanceMetric) {
    DeviceTile<T>(GetDeviceKernelId(), this->mParametersNumber, apMatrixA, aRowsNumber, aColumnsNumber, aRowOffset,
                  aColumnOffset, aLocation1, aLocation2, apLocalTheta, aDistanceMetric);
}

template class UnivariateMaternStationary<float>;
template class UnivariateMaternStationary<double>;
template class BivariateMaternParsimonious<float>;
template class BivariateMaternParsimonious<double>;

}  // namespace kernels

// ---------------------------------------------------------------- CSV data files
namespace dataLoader {
namespace csv {

template<typename T>
CSVLoader<T> *CSVLoader<T>::GetInstance() {
    static CSVLoader<T> instance;
    return &instance;
}

template<typename T>
void CSVLoader<T>::ReadData(configurations::Configurations &aConfigurations, std::vector<T> &aMeasurementsMatrix,
                            std::vector<T> &aXLocations, std::vector<T> &aYLocations, std::vector<T> &aZLocations,
                            const int &aP) {
    if (aConfigurations.GetDataPath().empty()) throw std::runtime_error("Please enter the path to data file.");
    const std::string path = aConfigurations.GetDataPath();
    const bool two_d = aConfigurations.GetDimension() == common::Dimension2D;
    std::ifstream file(path);
    if (!file.is_open()) throw std::runtime_error("Cannot read locations file: " + path);
    std::string line;
    int lines = 0;
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::vector<double> v;
        std::istringstream iss(line);
        std::string tok;
        while (std::getline(iss, tok, ',')) v.push_back(std::stod(tok));
        const size_t coords = two_d ? 2 : 3;
        if (v.size() < coords + 1)
            throw std::runtime_error("The data dimensions in the provided file isn't consistent with the dimensions input.");
        if (v.size() < coords + aP)
            throw std::runtime_error("The data P in the provided file isn't consistent with the Kernel's P.");
        aXLocations.push_back(static_cast<T>(v[0]));
        aYLocations.push_back(static_cast<T>(v[1]));
        if (!two_d) aZLocations.push_back(static_cast<T>(v[2]));
        for (int q = 0; q < aP; ++q) aMeasurementsMatrix.push_back(static_cast<T>(v[coords + q]));
        ++lines;
    }
    aConfigurations.SetProblemSize(lines * aP / aConfigurations.GetTimeSlot());
}

template<typename T>
void CSVLoader<T>::WriteData(const T &aMatrixPointer, const int &aProblemSize, const int &aP,
                             std::string &aLoggerPath, dataunits::Locations<T> &aLocations) {
    if (aLoggerPath.empty()) aLoggerPath = "synthetic_ds";
    else aLoggerPath += (aLoggerPath.back() == '/') ? "synthetic_ds" : "/synthetic_ds";
    std::error_code ec;
    std::filesystem::create_directories(aLoggerPath, ec);
    if (ec || !std::filesystem::exists(aLoggerPath)) throw std::runtime_error("Error creating directory: " + aLoggerPath);
    const int nloc = aProblemSize / aP;
    size_t i = 1;
    const std::string log_base = aLoggerPath + "/log_" + std::to_string(nloc) + "_";
    while (std::filesystem::exists(log_base + std::to_string(i))) ++i;
    std::ofstream out(aLoggerPath + "/SYN_" + std::to_string(nloc) + "_" + std::to_string(i));
    const T *z = &aMatrixPointer;
    out << std::setprecision(15);
    for (int l = 0; l < nloc; ++l) {
        out << aLocations.GetLocationX()[l] << ',' << aLocations.GetLocationY()[l];
        if (aLocations.GetLocationZ()) out << ',' << aLocations.GetLocationZ()[l];
        for (int q = 0; q < aP; ++q) out << ',' << z[static_cast<size_t>(l) * aP + q];
        out << '\n';
    }
}

template class CSVLoader<float>;
template class CSVLoader<double>;

}  // namespace csv
}  // namespace dataLoader

// ---------------------------------------------------------------- results
namespace results {

Results *Results::GetInstance() {
    static Results instance;
    return &instance;
}

namespace {
std::string Num(double v) { return std::to_string(v); }
std::string Vec(const std::vector<double> &v) {
    std::ostringstream oss;
    oss << "[ ";
    for (double x : v) oss << std::fixed << std::setprecision(8) << x << " ";
    oss << "]";
    return oss.str();
}
}  // namespace

void Results::SetIsSynthetic(bool a, const std::string &k) { UpdateDictionary(k.empty() ? "Data is Synthetic" : k, a ? "Yes" : "No"); }
void Results::SetGeneratedLocationsNumber(int a, const std::string &k) {
    UpdateDictionary(k.empty() ? "Number of Locations" : k, std::to_string(a));
}
void Results::SetTotalDataGenerationExecutionTime(double a, const std::string &k) {
    UpdateDictionary(k.empty() ? "Total Data Generation Execution Time" : k, Num(a));
}
void Results::SetTotalDataGenerationFlops(double a, const std::string &k) {
    UpdateDictionary(k.empty() ? "Total Data Generation Gflop/s" : k, Num(a));
}
void Results::SetLogLikValue(double a, const std::string &k) {
    mLogLikValue = a;
    UpdateDictionary(k.empty() ? "Final Log Likelihood Value" : k, Num(a));
}
void Results::SetMLEIterations(int a, const std::string &k) {
    mMLEIterations = a;
    UpdateDictionary(k.empty() ? "Number of MLE Iterations" : k, std::to_string(a));
}
void Results::SetMaximumTheta(const std::vector<double> &a, const std::string &k) {
    mMaximumTheta = a;
    UpdateDictionary(k.empty() ? "Found Maximum Theta at" : k, Vec(a));
}
void Results::SetTotalModelingExecutionTime(double a, const std::string &k) {
    mTotalModelingTime = a;
    UpdateDictionary(k.empty() ? "Total Modeling Execution Time" : k, Num(a));
    if (mMLEIterations > 0) UpdateDictionary("Average Time Modeling per Iteration", Num(a / mMLEIterations));
}
void Results::SetTotalModelingFlops(double a, const std::string &k) {
    mTotalModelingFlops = a;
    UpdateDictionary(k.empty() ? "Total Modeling Flops" : k, Num(a));
    if (mMLEIterations > 0) UpdateDictionary("Average Flops per Iteration", Num(a / mMLEIterations));
}
double Results::GetAverageModelingExecutionTime() const {
    return mMLEIterations > 0 ? mTotalModelingTime / mMLEIterations : 0.0;
}
double Results::GetAverageModelingFlops() const {
    return mMLEIterations > 0 ? mTotalModelingFlops / mMLEIterations : 0.0;
}
void Results::SetZMiss(int a, const std::string &k) {
    UpdateDictionary(k.empty() ? "Number of Missing Observations" : k, std::to_string(a));
}
void Results::SetMSPEError(double a, const std::string &k) {
    mMSPEError = a;
    UpdateDictionary(k.empty() ? "Mean Square Error MSPE" : k, Num(a));
}
void Results::SetMSPEExecutionTime(double a, const std::string &k) {
    UpdateDictionary(k.empty() ? "MSPE Prediction Execution Time" : k, Num(a));
}
void Results::SetMSPEFlops(double a, const std::string &k) { UpdateDictionary(k.empty() ? "MSPE Gflop/s" : k, Num(a)); }
void Results::SetPredictedMissedValues(std::vector<double> a, const std::string &k) {
    mPredicted = a;
    UpdateDictionary(k.empty() ? "Predicted Values" : k, Vec(a));
}
void Results::PrintEndSummary() const {
    std::cout << "********************SUMMARY**********************" << std::endl;
    for (const auto &e : mSummaryDictionary) std::cout << "#" << e.first << ": " << e.second << std::endl;
    std::cout << "*************************************************" << std::endl;
}

}  // namespace results
}  // namespace exageostat

// ---------------------------------------------------------------- hardware
ExaGeoStatHardware::ExaGeoStatHardware(const exageostat::common::Computation &aComputation, const int &,
                                       const int &aGpuNumber) {
    if (aComputation != exageostat::common::EXACT_DENSE)
        throw std::domain_error("only EXACT_DENSE computation runs on the MI355X path");
    if (aGpuNumber > 1) throw std::domain_error("the MI355X path runs on one GPU");
    if (!exageostat::g_context) {
        exageostat::CheckStatus(egs_init(0, &exageostat::g_context), "ExaGeoStatHardware");
        mOwner = true;
    }
}

ExaGeoStatHardware::ExaGeoStatHardware(ExaGeoStatHardware &&aOther) noexcept : mOwner(aOther.mOwner) {
    aOther.mOwner = false;
}

ExaGeoStatHardware::~ExaGeoStatHardware() {
    if (mOwner && exageostat::g_context) {
        egs_finalize(exageostat::g_context);
        exageostat::g_context = nullptr;
    }
}

egs_context_t ExaGeoStatHardware::GetContext() {
    if (!exageostat::g_context) throw std::runtime_error("ExaGeoStatHardware is not initialised");
    return exageostat::g_context;
}

// ---------------------------------------------------------------- data
template<typename T>
egs_problem_t ExaGeoStatData<T>::GetProblem(egs_context_t aContext, int aKernelId, int aMetric, int aDts) {
    if (mProblem && mKernelId == aKernelId && mMetric == aMetric && mDts == aDts) return mProblem;
    ReleaseProblem();
    auto *l = mLocations.get();
    const size_t n = l->GetSize();
    auto x = exageostat::ToDouble(l->GetLocationX(), n), y = exageostat::ToDouble(l->GetLocationY(), n);
    auto z = exageostat::ToDouble(l->GetLocationZ(), n);
    exageostat::CheckStatus(egs_problem_create(aContext, aKernelId, (int64_t) n, x.data(), y.data(),
                                               z.empty() ? nullptr : z.data(), aMetric, aDts, &mProblem),
                            "egs_problem_create");
    mKernelId = aKernelId;
    mMetric = aMetric;
    mDts = aDts;
    mPlan[0] = EGS_PRECISION_FP64;   // a new problem starts under the FP64 plan
    return mProblem;
}

template<typename T>
void ExaGeoStatData<T>::ReleaseProblem() {
    if (mProblem) egs_problem_destroy(mProblem);
    mProblem = nullptr;
    mKernelId = mMetric = mDts = -1;
    mPlan[0] = mPlan[1] = mPlan[2] = -1;
}

template<typename T>
void ExaGeoStatData<T>::SetProblemPlan(int aMode, int aBandDP, int aBandSP) {
    if (!mProblem) throw std::runtime_error("SetProblemPlan: no device problem");
    const bool fp64 = aMode == EGS_PRECISION_FP64;
    if (mPlan[0] == aMode && (fp64 || (mPlan[1] == aBandDP && mPlan[2] == aBandSP))) return;
    exageostat::CheckStatus(egs_problem_set_precision(mProblem, aMode, aBandDP, aBandSP), "egs_problem_set_precision");
    mPlan[0] = aMode;
    mPlan[1] = aBandDP;
    mPlan[2] = aBandSP;
}

template class ExaGeoStatData<float>;
template class ExaGeoStatData<double>;

// ---------------------------------------------------------------- API
namespace exageostat {

namespace {

int PrecisionMode(configurations::Configurations &aConf) {
    switch (aConf.GetPrecision()) {
        case common::DOUBLE: return EGS_PRECISION_FP64;
        case common::MIXED: return EGS_PRECISION_MIXED;
        default: throw std::domain_error("SINGLE precision is not available; use DOUBLE or MIXED");
    }
}

template<typename T>
dataunits::VectorDescriptor<T> &AsVector(void *apDescriptor, const char *aWhere) {
    if (!apDescriptor) throw std::runtime_error(std::string(aWhere) + ": null vector descriptor");
    return *static_cast<dataunits::VectorDescriptor<T> *>(apDescriptor);
}

egs_problem_t AsMatrix(void *apDescriptor, const char *aWhere) {
    if (!apDescriptor) throw std::runtime_error(std::string(aWhere) + ": null matrix descriptor");
    return static_cast<egs_problem_t>(apDescriptor);
}

// Fisher-Yates shuffle with rand() as PredictionHelpers<T>::Shuffle
// (PredictionHelpers.cpp:141-240): value arrays and locations move together.
template<typename T>
void ShuffleTogether(std::vector<std::vector<double>> &aParts, dataunits::Locations<T> &aLoc, int aSize) {
    if (aSize <= 1) return;
    const bool has_z = aLoc.GetDimension() != common::Dimension2D;
    for (size_t i = 0; i + 1 < static_cast<size_t>(aSize); ++i) {
        const size_t j = i + std::rand() / (RAND_MAX / (aSize - i) + 1);
        for (auto &part : aParts) std::swap(part[j], part[i]);
        std::swap(aLoc.GetLocationX()[j], aLoc.GetLocationX()[i]);
        std::swap(aLoc.GetLocationY()[j], aLoc.GetLocationY()[i]);
        if (has_z) std::swap(aLoc.GetLocationZ()[j], aLoc.GetLocationZ()[i]);
    }
}

}  // namespace

namespace linearAlgebra {
namespace dense {

template<typename T>
T HipDense<T>::ExaGeoStatMLETile(std::unique_ptr<ExaGeoStatData<T>> &aData,
                                 configurations::Configurations &aConfigurations, const double *apTheta,
                                 T *apMeasurementsMatrix, const kernels::Kernel<T> &aKernel) {
    if (!aData) throw std::runtime_error("ExaGeoStatMLETile: no data");
    if (!apTheta) throw std::runtime_error("ExaGeoStatMLETile: theta is null");
    const size_t nz = static_cast<size_t>(aData->GetLocations()->GetSize()) * aKernel.GetVariablesNumber();
    egs_problem_t p = aData->GetProblem(ExaGeoStatHardware::GetContext(), aKernel.GetDeviceKernelId(),
                                        aConfigurations.GetDistanceMetric(), aConfigurations.GetDenseTileSize());
    // Z restored from the measurements before every evaluation
    // (ChameleonImplementation.cpp:113-125); the device copy is never overwritten
    const std::vector<double> z = apMeasurementsMatrix ? ToDouble(apMeasurementsMatrix, nz) : aData->GetMeasurements();
    if (z.size() != nz) throw std::runtime_error("ExaGeoStatMLETile: no measurements");
    CheckStatus(egs_problem_set_observations(p, z.data(), (int64_t) nz), "egs_problem_set_observations");
    aData->SetProblemPlan(PrecisionMode(aConfigurations), aConfigurations.GetDenseBandDP(),
                          aConfigurations.GetBandDenseSP());
    double llh = 0.0;
    CheckStatus(egs_loglik(p, apTheta, &llh, nullptr, nullptr), "ExaGeoStatMLETile");
    double t_cov = 0.0, t_fact = 0.0, t_solve = 0.0;
    egs_last_timings(p, &t_cov, &t_fact, &t_solve);
    const int np = aKernel.GetParametersNumbers();
    std::vector<double> theta(apTheta, apTheta + np);
    aData->SetMleIterations(aData->GetMleIterations() + 1);
    auto *res = results::Results::GetInstance();
    res->SetMLEIterations(aData->GetMleIterations());
    res->SetMaximumTheta(theta);
    res->SetLogLikValue(llh);
    // the per-evaluation lines of ChameleonImplementation.cpp:273-310
    if (Verbose()) {
        std::printf("\t%d- Model Parameters (", aData->GetMleIterations());
        for (int i = 0; i < np; ++i) std::printf("%.8f%s", theta[i], i + 1 < np ? ", " : "");
        std::printf(")----> LogLi: %.18f\n", llh);
    }
    if (configurations::Configurations::GetVerbosity() == common::DETAILED_MODE) {
        const double n = static_cast<double>(nz);
        std::printf("---- Facto Time: %g\n---- dtrsm Time: %g\n---- Matrix Generation Time: %g\n"
                    "---- Total Time: %g\n---- Gflop/s: %g\n",
                    t_fact * 1e-3, t_solve * 1e-3, t_cov * 1e-3, (t_fact + t_solve) * 1e-3,
                    (n * n * n / 3.0 + 2.0 * n * n) / 1e9 / std::max((t_fact + t_solve) * 1e-3, 1e-12));
    }
    return static_cast<T>(llh);
}

template<typename T>
void HipDense<T>::ExaGeoStatLapackCopyTile(const common::UpperLower &, void *apA, void *apB) {
    // the Z -> Z_COPY copy of ChameleonImplementation.cpp:329-335 on vector descriptors
    auto &a = AsVector<T>(apA, "ExaGeoStatLapackCopyTile");
    auto &b = AsVector<T>(apB, "ExaGeoStatLapackCopyTile");
    if (a.size != b.size) throw std::runtime_error("ExaGeoStatLapackCopyTile: descriptor sizes differ");
    std::copy(a.data, a.data + a.size, b.data);
}

template<typename T>
void HipDense<T>::ExaGeoStatSequenceWait(void *apSequence) {
    if (!apSequence) throw std::runtime_error("ExaGeoStatSequenceWait: null sequence");
    CheckStatus(egs_synchronize(static_cast<egs_context_t>(apSequence)), "ExaGeoStatSequenceWait");
}

template<typename T>
void HipDense<T>::ExaGeoStatCreateSequence(void *apSequence) {
    // the reference passes the address of its sequence pointer
    if (!apSequence) throw std::runtime_error("ExaGeoStatCreateSequence: null sequence");
    *static_cast<egs_context_t *>(apSequence) = ExaGeoStatHardware::GetContext();
}

template<typename T>
void HipDense<T>::ExaGeoStatPotrfTile(const common::UpperLower &aUpperLower, void *apA, int, void *, void *,
                                      const int &, const int &) {
    if (aUpperLower != common::EXAGEOSTAT_LOWER)
        throw std::domain_error("ExaGeoStatPotrfTile: the tiles hold the lower triangle");
    CheckStatus(egs_potrf(AsMatrix(apA, "ExaGeoStatPotrfTile")), "ExaGeoStatPotrfTile");
}

template<typename T>
void HipDense<T>::ExaGeoStatTrsmTile(const common::Side &aSide, const common::UpperLower &aUpperLower,
                                     const common::Trans &aTrans, const common::Diag &aDiag, const T &aAlpha,
                                     void *apA, void *, void *, void *apZ, const int &) {
    if (aSide != common::EXAGEOSTAT_LEFT || aUpperLower != common::EXAGEOSTAT_LOWER ||
        aDiag != common::EXAGEOSTAT_NON_UNIT || aTrans == common::EXAGEOSTAT_CONJ_TRANS)
        throw std::domain_error("ExaGeoStatTrsmTile: the factor is lower, non-unit, applied from the left");
    auto &z = AsVector<T>(apZ, "ExaGeoStatTrsmTile");
    std::vector<double> v = ToDouble(z.data, static_cast<size_t>(z.size));
    CheckStatus(egs_trsm_vector(AsMatrix(apA, "ExaGeoStatTrsmTile"), aTrans == common::EXAGEOSTAT_TRANS ? 1 : 0,
                                static_cast<double>(aAlpha), v.data(), z.size),
                "ExaGeoStatTrsmTile");
    for (int64_t i = 0; i < z.size; ++i) z.data[i] = static_cast<T>(v[i]);
}

template<typename T>
void HipDense<T>::GenerateObservationsVector(configurations::Configurations &aConfigurations,
                                             std::unique_ptr<ExaGeoStatData<T>> &aData,
                                             dataunits::Locations<T> *apLocation1, dataunits::Locations<T> *,
                                             dataunits::Locations<T> *, const int &aDistanceMetric,
                                             const kernels::Kernel<T> &aKernel) {
    // LinearAlgebraMethods.cpp:333-449: z = L(initial theta) e, e = dlarnv(3, {seed, seed, seed, 1})
    if (!aData) throw std::runtime_error("GenerateObservationsVector: no data");
    if (apLocation1 && apLocation1 != aData->GetLocations())
        throw std::domain_error("GenerateObservationsVector: observations are generated at the data's locations");
    std::vector<double> theta = aConfigurations.GetInitialTheta();
    if (static_cast<int>(theta.size()) < aKernel.GetParametersNumbers())
        throw std::domain_error("initial theta needs " + std::to_string(aKernel.GetParametersNumbers()) + " values");
    egs_problem_t p = aData->GetProblem(ExaGeoStatHardware::GetContext(), aKernel.GetDeviceKernelId(),
                                        aDistanceMetric, aConfigurations.GetDenseTileSize());
    aData->SetProblemPlan(EGS_PRECISION_FP64, 0, 0);   // data is generated with the FP64 factor
    std::vector<double> &obs = aData->GetMeasurements();
    obs.assign(static_cast<size_t>(aData->GetLocations()->GetSize()) * aKernel.GetVariablesNumber(), 0.0);
    CheckStatus(egs_generate_observations(p, theta.data(), aConfigurations.GetSeed(), obs.data()),
                "GenerateObservationsVector");
}

template<typename T>
void HipDense<T>::ExaGeoStatPosvTile(const common::UpperLower &aUpperLower, void *apA, void *apB) {
    // dposv (LinearAlgebraMethods.cpp:1421-1428): factor Sigma, then L L^T x = b
    ExaGeoStatPotrfTile(aUpperLower, apA, 0, nullptr, nullptr, 0, 0);
    const T one = 1;
    ExaGeoStatTrsmTile(common::EXAGEOSTAT_LEFT, aUpperLower, common::EXAGEOSTAT_NO_TRANS, common::EXAGEOSTAT_NON_UNIT,
                       one, apA, nullptr, nullptr, apB, 0);
    ExaGeoStatTrsmTile(common::EXAGEOSTAT_LEFT, aUpperLower, common::EXAGEOSTAT_TRANS, common::EXAGEOSTAT_NON_UNIT,
                       one, apA, nullptr, nullptr, apB, 0);
}

template<typename T>
T *HipDense<T>::ExaGeoStatMLEPredictTile(std::unique_ptr<ExaGeoStatData<T>> &, T *apTheta, const int &aZMissNumber,
                                         const int &aZObsNumber, T *apZObs, T *apZActual, T *apZMiss,
                                         configurations::Configurations &aConfiguration,
                                         dataunits::Locations<T> &aMissLocations,
                                         dataunits::Locations<T> &aObsLocations, const kernels::Kernel<T> &aKernel) {
    // LinearAlgebraMethods.cpp:451-605: C22 = Sigma(obs, obs) factored on the
    // GPU, zmiss = C12 C22^{-1} zobs with C12 regenerated, the three errors
    if (!apTheta || !apZObs || !apZMiss || aZMissNumber <= 0 || aZObsNumber <= 0)
        throw std::runtime_error("ExaGeoStatMLEPredictTile: invalid arguments");
    const int p = aKernel.GetVariablesNumber();
    const bool has_z = aObsLocations.GetDimension() != common::Dimension2D;
    const double t0 = Seconds();
    auto ox = ToDouble(aObsLocations.GetLocationX(), aZObsNumber), oy = ToDouble(aObsLocations.GetLocationY(), aZObsNumber);
    auto oz = ToDouble(aObsLocations.GetLocationZ(), has_z ? aZObsNumber : 0);
    auto mx = ToDouble(aMissLocations.GetLocationX(), aZMissNumber), my = ToDouble(aMissLocations.GetLocationY(), aZMissNumber);
    auto mz = ToDouble(aMissLocations.GetLocationZ(), has_z ? aZMissNumber : 0);
    egs_problem_t op = nullptr;
    CheckStatus(egs_problem_create(ExaGeoStatHardware::GetContext(), aKernel.GetDeviceKernelId(), aZObsNumber,
                                   ox.data(), oy.data(), oz.empty() ? nullptr : oz.data(),
                                   aConfiguration.GetDistanceMetric(), aConfiguration.GetDenseTileSize(), &op),
                "egs_problem_create");
    std::unique_ptr<egs_problem_s, int (*)(egs_problem_t)> guard(op, egs_problem_destroy);
    auto zobs = ToDouble(apZObs, static_cast<size_t>(aZObsNumber) * p);
    CheckStatus(egs_problem_set_observations(op, zobs.data(), (int64_t) zobs.size()), "egs_problem_set_observations");
    if (PrecisionMode(aConfiguration) != EGS_PRECISION_FP64)
        CheckStatus(egs_problem_set_precision(op, PrecisionMode(aConfiguration), aConfiguration.GetDenseBandDP(),
                                              aConfiguration.GetBandDenseSP()),
                    "egs_problem_set_precision");
    std::vector<double> theta(apTheta, apTheta + aKernel.GetParametersNumbers());
    std::vector<double> zmiss(static_cast<size_t>(aZMissNumber) * p);
    auto zact = ToDouble(apZActual, apZActual ? static_cast<size_t>(aZMissNumber) * p : 0);
    double m3[3] = {-1.0, 0.0, 0.0};
    CheckStatus(egs_predict_mspe(op, theta.data(), aZMissNumber, mx.data(), my.data(), mz.empty() ? nullptr : mz.data(),
                                 zmiss.data(), zact.empty() ? nullptr : zact.data(), m3),
                "ExaGeoStatMLEPredictTile");
    for (size_t i = 0; i < zmiss.size(); ++i) apZMiss[i] = static_cast<T>(zmiss[i]);
    const double dt = Seconds() - t0;
    auto *res = results::Results::GetInstance();
    res->SetMSPEExecutionTime(dt);
    const double no = static_cast<double>(aZObsNumber) * p;
    res->SetMSPEFlops((no * no * no / 3.0 + 2.0 * no * aZMissNumber * p) / 1e9 / std::max(dt, 1e-12));
    if (apZActual) res->SetMSPEError(m3[0]);
    T *all = new T[3];
    for (int i = 0; i < 3; ++i) all[i] = static_cast<T>(m3[i]);
    return all;
}

template class HipDense<float>;
template class HipDense<double>;

}  // namespace dense

template<typename T>
std::unique_ptr<LinearAlgebraMethods<T>> LinearAlgebraFactory<T>::CreateLinearAlgebraSolver(
    common::Computation aComputation) {
    if (aComputation == common::EXACT_DENSE) return std::make_unique<dense::HipDense<T>>();
    if (aComputation == common::TILE_LOW_RANK)
        throw std::runtime_error("Tile low rank generation isn't supported on the MI355X path (no HiCMA)");
    throw std::domain_error("only EXACT_DENSE computation runs on the MI355X path");
}

template class LinearAlgebraFactory<float>;
template class LinearAlgebraFactory<double>;

}  // namespace linearAlgebra

namespace api {

template<typename T>
void ExaGeoStat<T>::ExaGeoStatLoadData(configurations::Configurations &aConfigurations,
                                       std::unique_ptr<ExaGeoStatData<T>> &aData) {
    if (aConfigurations.GetIsSynthetic()) {
        ExaGeoStatGenerateData(aConfigurations, aData);
        return;
    }
    std::srand(aConfigurations.GetSeed());
    auto kernel = CreateKernel<T>(aConfigurations);
    const Dimension dim = aConfigurations.GetDimension();
    if (dim == common::DimensionST) throw std::domain_error("space-time data is not on the MI355X path");
    // real data from a CSV file (CSVLoader<T>::LoadData)
    std::vector<T> z, xs, ys, zs;
    dataLoader::csv::CSVLoader<T>::GetInstance()->ReadData(aConfigurations, z, xs, ys, zs,
                                                           kernel->GetVariablesNumber());
    const int nloc = static_cast<int>(xs.size());
    aData = std::make_unique<ExaGeoStatData<T>>(nloc, dim);
    aData->GetLocations()->SetLocationX(*xs.data(), nloc);
    aData->GetLocations()->SetLocationY(*ys.data(), nloc);
    if (!zs.empty()) aData->GetLocations()->SetLocationZ(*zs.data(), nloc);
    aData->GetMeasurements() = ToDouble(z);
    results::Results::GetInstance()->SetIsSynthetic(false);
    results::Results::GetInstance()->SetGeneratedLocationsNumber(nloc);
}

template<typename T>
void ExaGeoStat<T>::ExaGeoStatGenerateData(configurations::Configurations &aConfigurations,
                                           std::unique_ptr<ExaGeoStatData<T>> &aData) {
    std::srand(aConfigurations.GetSeed());
    auto kernel = CreateKernel<T>(aConfigurations);
    const Dimension dim = aConfigurations.GetDimension();
    if (dim == common::DimensionST) throw std::domain_error("space-time data is not on the MI355X path");
    auto solver = linearAlgebra::LinearAlgebraFactory<T>::CreateLinearAlgebraSolver(aConfigurations.GetComputation());
    const int n = aConfigurations.GetProblemSize();
    const double t0 = Seconds();
    std::vector<double> x(n), y(n), z(dim == common::Dimension3D ? n : 0);
    CheckStatus(egs_generate_locations(n, dim == common::Dimension3D ? 3 : 2, x.data(), y.data(),
                                       z.empty() ? nullptr : z.data()),
                "GenerateLocations");
    aData = std::make_unique<ExaGeoStatData<T>>(n, dim);
    auto *loc = aData->GetLocations();
    for (int i = 0; i < n; ++i) {
        loc->GetLocationX()[i] = static_cast<T>(x[i]);
        loc->GetLocationY()[i] = static_cast<T>(y[i]);
        if (!z.empty()) loc->GetLocationZ()[i] = static_cast<T>(z[i]);
    }
    solver->GenerateObservationsVector(aConfigurations, aData, loc, loc, nullptr, aConfigurations.GetDistanceMetric(),
                                       *kernel);
    std::vector<double> &obs = aData->GetMeasurements();
    if (aConfigurations.GetLogger()) {
        std::vector<T> zt(obs.begin(), obs.end());
        std::string path = aConfigurations.GetLoggerPath();
        dataLoader::csv::CSVLoader<T>::GetInstance()->WriteData(*zt.data(), static_cast<int>(zt.size()),
                                                                kernel->GetVariablesNumber(), path, *loc);
    }
    const double t1 = Seconds();
    const double nn = static_cast<double>(obs.size());
    auto *res = results::Results::GetInstance();
    res->SetIsSynthetic(true);
    res->SetGeneratedLocationsNumber(n);
    res->SetTotalDataGenerationExecutionTime(t1 - t0);
    res->SetTotalDataGenerationFlops((nn * nn * nn / 3.0 + nn * nn) / 1e9 / std::max(t1 - t0, 1e-12));
}

template<typename T>
T ExaGeoStat<T>::ExaGeoStatDataModeling(configurations::Configurations &aConfigurations,
                                        std::unique_ptr<ExaGeoStatData<T>> &aData, T *apMeasurementsMatrix) {
    if (!aData) throw std::runtime_error("ExaGeoStatDataModeling: no data");
    auto kernel = CreateKernel<T>(aConfigurations);
    auto solver = linearAlgebra::LinearAlgebraFactory<T>::CreateLinearAlgebraSolver(aConfigurations.GetComputation());
    aConfigurations.InitializeAllTheta();
    const int np = kernel->GetParametersNumbers();
    const size_t nz = static_cast<size_t>(aData->GetLocations()->GetSize()) * kernel->GetVariablesNumber();
    std::vector<double> &obs = aData->GetMeasurements();
    if (apMeasurementsMatrix) obs = ToDouble(apMeasurementsMatrix, nz);
    if (obs.size() != nz) throw std::runtime_error("ExaGeoStatDataModeling: no measurements");

    std::vector<double> lb = aConfigurations.GetLowerBounds(), ub = aConfigurations.GetUpperBounds();
    std::vector<double> x = aConfigurations.GetStartingTheta();
    std::vector<int> free_idx;
    for (int i = 0; i < np; ++i) {
        if (ub[i] > lb[i]) free_idx.push_back(i);
        else if (ub[i] < lb[i]) throw std::domain_error("upper bound below lower bound for parameter " + std::to_string(i));
        x[i] = std::min(ub[i], std::max(lb[i], x[i]));
    }

    // StarpuRuntimeSolver<T>::DataModelingAPI (StarpuRuntimeSolver.cpp:67-79):
    // one ExaGeoStatMLETile per optimizer point, with the modeling totals
    struct State {
        linearAlgebra::LinearAlgebraMethods<T> *solver;
        std::unique_ptr<ExaGeoStatData<T>> *data;
        configurations::Configurations *conf;
        const kernels::Kernel<T> *kernel;
        std::vector<double> theta;
        const std::vector<int> *free_idx;
        double n, total_time = 0.0, total_flops = 0.0, best = -INFINITY;
        std::vector<double> best_theta;
        bool failed = false;
        std::string error;
    } st{solver.get(), &aData, &aConfigurations, kernel.get(), x, &free_idx, static_cast<double>(nz)};
    auto objective = [](const double *v, void *user) -> double {
        auto *s = static_cast<State *>(user);
        for (size_t j = 0; j < s->free_idx->size(); ++j) s->theta[(*s->free_idx)[j]] = v[j];
        double llh = 0.0;
        const double t0 = Seconds();
        try {
            // the observations are the data's measurements (already set)
            llh = static_cast<double>(s->solver->ExaGeoStatMLETile(*s->data, *s->conf, s->theta.data(), nullptr,
                                                                    *s->kernel));
        } catch (const std::exception &e) {
            // the optimizer is C code: record the failure and stop it with NaN
            if (!s->failed) {
                s->failed = true;
                s->error = e.what();
            }
            return NAN;
        }
        s->total_time += Seconds() - t0;
        s->total_flops += (s->n * s->n * s->n / 3.0 + 2.0 * s->n * s->n) / 1e9;
        auto *res = results::Results::GetInstance();
        res->SetTotalModelingExecutionTime(s->total_time);
        res->SetTotalModelingFlops(s->total_flops);
        if (llh > s->best) {
            s->best = llh;
            s->best_theta = s->theta;
        }
        return -llh;
    };
    const int nf = static_cast<int>(free_idx.size());
    int used = 0;
    auto search = [&](const std::vector<double> &start, int budget) {
        std::vector<double> xf(nf), lf(nf), uf(nf);
        for (int j = 0; j < nf; ++j) {
            xf[j] = start[free_idx[j]];
            lf[j] = lb[free_idx[j]];
            uf[j] = ub[free_idx[j]];
        }
        double fmin = 0.0;
        int nev = 0, reason = 0;
        egs_bobyqa(nf, xf.data(), lf.data(), uf.data(), aConfigurations.GetTolerance(), budget, 0.0, objective, &st,
                   &fmin, &nev, &reason);
        used += nev;
    };
    if (free_idx.empty()) objective(nullptr, &st);
    else search(x, aConfigurations.GetMaxMleIterations());
    // Under MIXED precision the search ends with an FP64 polish from the mixed
    // optimum (as egs_data_modeling does): the reported optimum is FP64's
    configurations::Configurations c64 = aConfigurations;
    if (!st.failed && nf > 0 && PrecisionMode(aConfigurations) != EGS_PRECISION_FP64 &&
        used < aConfigurations.GetMaxMleIterations()) {
        c64.SetPrecision(common::DOUBLE);
        st.conf = &c64;
        const std::vector<double> start = st.best_theta;
        st.theta = start;
        st.best = -INFINITY;
        search(start, aConfigurations.GetMaxMleIterations() - used);
        st.conf = &aConfigurations;
    }
    if (st.failed) throw std::runtime_error(st.error);
    aConfigurations.SetStartingTheta(st.best_theta);
    aConfigurations.SetEstimatedTheta(st.best_theta);
    if (Verbose()) {
        std::printf("--> Final Theta Values (");
        for (size_t i = 0; i < st.best_theta.size(); ++i)
            std::printf("%.8f%s", st.best_theta[i], i + 1 < st.best_theta.size() ? ", " : "");
        std::printf(")\n");
    }
    return static_cast<T>(st.best);
}

template<typename T>
void ExaGeoStat<T>::ExaGeoStatPrediction(configurations::Configurations &aConfigurations,
                                         std::unique_ptr<ExaGeoStatData<T>> &aData, T *apMeasurementsMatrix,
                                         dataunits::Locations<T> *apTrainLocations,
                                         dataunits::Locations<T> *apTestLocations) {
    if (!aData) throw std::runtime_error("ExaGeoStatPrediction: no data");
    auto kernel = CreateKernel<T>(aConfigurations);
    aConfigurations.InitializeAllTheta();
    const int np = kernel->GetParametersNumbers();
    bool can_predict = true;
    for (int i = 0; i < np; ++i)
        if (aConfigurations.GetEstimatedTheta()[i] == -1) can_predict = false;
    if (!can_predict &&
        (aConfigurations.GetIsMLOEMMOM() || aConfigurations.GetIsMSPE() || aConfigurations.GetIsFisher()))
        throw std::runtime_error("Can't predict without an estimated theta, please either pass --etheta or run the "
                                 "modeling module before prediction");
    if (aConfigurations.GetIsIDW() || aConfigurations.GetIsMLOEMMOM() || aConfigurations.GetIsFisher())
        throw std::domain_error("IDW, MLOE-MMOM and Fisher are not on the MI355X likelihood path");
    if (aConfigurations.GetIsNonGaussian()) throw std::domain_error("the non-Gaussian model is not on this path");

    const int p = kernel->GetVariablesNumber();
    const bool explicit_sets = apTrainLocations && apTestLocations;
    int z_miss = explicit_sets ? apTestLocations->GetSize() : aConfigurations.GetUnknownObservationsNb();
    int n_obs = explicit_sets ? apTrainLocations->GetSize() : aConfigurations.CalculateZObsNumber();
    aConfigurations.SetObservationNumber(n_obs);
    if (z_miss <= 0) return;
    results::Results::GetInstance()->SetZMiss(z_miss);
    aConfigurations.SetUnknownObservationsNb(z_miss);

    const Dimension dim = aConfigurations.GetDimension();
    dataunits::Locations<T> miss(z_miss, dim), obs_loc(n_obs, dim);
    std::vector<double> z_obs, z_actual;
    if (!explicit_sets) {
        // PredictionHelpers<T>::PickRandomPoints (PredictionHelpers.cpp:23-138)
        const int n = aConfigurations.GetProblemSize();
        std::vector<double> z;
        if (apMeasurementsMatrix) z = ToDouble(apMeasurementsMatrix, static_cast<size_t>(n) * p);
        else z = aData->GetMeasurements();
        if (static_cast<int>(z.size()) < n * p) throw std::runtime_error("ExaGeoStatPrediction: no measurements");
        std::vector<std::vector<double>> parts(p, std::vector<double>(n));
        for (int v = 0; v < p; ++v)
            for (int i = 0; i < n; ++i) parts[v][i] = z[static_cast<size_t>(i) * p + v];
        dataunits::Locations<T> l(n, dim);
        auto *src = aData->GetLocations();
        for (int i = 0; i < n; ++i) {
            l.GetLocationX()[i] = src->GetLocationX()[i];
            l.GetLocationY()[i] = src->GetLocationY()[i];
            if (dim != common::Dimension2D) l.GetLocationZ()[i] = src->GetLocationZ()[i];
        }
        ShuffleTogether(parts, l, n);
        for (int i = 0; i < z_miss; ++i)
            for (int v = 0; v < p; ++v) z_actual.push_back(parts[v][i]);
        for (int i = 0; i < n_obs; ++i)
            for (int v = 0; v < p; ++v) z_obs.push_back(parts[v][z_miss + i]);
        for (int i = 0; i < z_miss; ++i) {
            miss.GetLocationX()[i] = l.GetLocationX()[i];
            miss.GetLocationY()[i] = l.GetLocationY()[i];
            if (dim != common::Dimension2D) miss.GetLocationZ()[i] = l.GetLocationZ()[i];
        }
        for (int i = 0; i < n_obs; ++i) {
            obs_loc.GetLocationX()[i] = l.GetLocationX()[z_miss + i];
            obs_loc.GetLocationY()[i] = l.GetLocationY()[z_miss + i];
            if (dim != common::Dimension2D) obs_loc.GetLocationZ()[i] = l.GetLocationZ()[z_miss + i];
        }
        // SortInplace (PredictionHelpers.cpp:299-311) leaves the order unchanged.
    } else {
        // With explicit train/test sets the reference sizes the observed system
        // by n - zmiss of the configuration (Configurations.cpp:364-366), so
        // only the first n - zmiss training points take part.
        n_obs = std::min(n_obs, aConfigurations.CalculateZObsNumber());
        if (n_obs <= 0) throw std::domain_error("ExaGeoStatPrediction: no observations left (n - zmiss <= 0)");
        if (!apMeasurementsMatrix) throw std::runtime_error("ExaGeoStatPrediction: measurements are required");
        obs_loc = dataunits::Locations<T>(n_obs, dim);
        for (int i = 0; i < n_obs; ++i) {
            obs_loc.GetLocationX()[i] = apTrainLocations->GetLocationX()[i];
            obs_loc.GetLocationY()[i] = apTrainLocations->GetLocationY()[i];
            if (dim != common::Dimension2D) obs_loc.GetLocationZ()[i] = apTrainLocations->GetLocationZ()[i];
        }
        for (int i = 0; i < z_miss; ++i) {
            miss.GetLocationX()[i] = apTestLocations->GetLocationX()[i];
            miss.GetLocationY()[i] = apTestLocations->GetLocationY()[i];
            if (dim != common::Dimension2D) miss.GetLocationZ()[i] = apTestLocations->GetLocationZ()[i];
        }
        z_obs = ToDouble(apMeasurementsMatrix, static_cast<size_t>(n_obs) * p);
    }

    // Prediction.cpp:165-197: the solver's ExaGeoStatMLEPredictTile
    auto solver = linearAlgebra::LinearAlgebraFactory<T>::CreateLinearAlgebraSolver(aConfigurations.GetComputation());
    std::vector<T> zobs_t(z_obs.begin(), z_obs.end()), zact_t(z_actual.begin(), z_actual.end());
    std::vector<T> zmiss_t(static_cast<size_t>(z_miss) * p);
    std::vector<T> theta(aConfigurations.GetEstimatedTheta().begin(), aConfigurations.GetEstimatedTheta().begin() + np);
    std::unique_ptr<T[]> errors(solver->ExaGeoStatMLEPredictTile(aData, theta.data(), z_miss, n_obs, zobs_t.data(),
                                                                 zact_t.empty() ? nullptr : zact_t.data(),
                                                                 zmiss_t.data(), aConfigurations, miss, obs_loc,
                                                                 *kernel));
    if (!z_actual.empty() && Verbose()) std::printf("\t\t- Prediction value: %f\n", static_cast<double>(errors[0]));
    results::Results::GetInstance()->SetPredictedMissedValues(std::vector<double>(zmiss_t.begin(), zmiss_t.end()));
}

template class ExaGeoStat<float>;
template class ExaGeoStat<double>;

}  // namespace api
}  // namespace exageostat
