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
    return mProblem;
}

template<typename T>
void ExaGeoStatData<T>::ReleaseProblem() {
    if (mProblem) egs_problem_destroy(mProblem);
    mProblem = nullptr;
    mKernelId = mMetric = mDts = -1;
}

template class ExaGeoStatData<float>;
template class ExaGeoStatData<double>;

// ---------------------------------------------------------------- API
namespace exageostat {
namespace api {

namespace {

int PrecisionMode(configurations::Configurations &aConf) {
    switch (aConf.GetPrecision()) {
        case common::DOUBLE: return EGS_PRECISION_FP64;
        case common::MIXED: return EGS_PRECISION_MIXED;
        default: throw std::domain_error("SINGLE precision is not available; use DOUBLE or MIXED");
    }
}

void ConfigurePrecision(egs_problem_t aProblem, configurations::Configurations &aConf) {
    const int mode = PrecisionMode(aConf);
    CheckStatus(egs_problem_set_precision(aProblem, mode, aConf.GetDenseBandDP(), aConf.GetBandDenseSP()),
                "egs_problem_set_precision");
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

template<typename T>
void ExaGeoStat<T>::ExaGeoStatLoadData(configurations::Configurations &aConfigurations,
                                       std::unique_ptr<ExaGeoStatData<T>> &aData) {
    std::srand(aConfigurations.GetSeed());
    auto kernel = CreateKernel<T>(aConfigurations);
    const Dimension dim = aConfigurations.GetDimension();
    if (dim == common::DimensionST) throw std::domain_error("space-time data is not on the MI355X path");
    if (!aConfigurations.GetIsSynthetic()) {
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
        return;
    }
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
    std::vector<double> theta = aConfigurations.GetInitialTheta();
    if (static_cast<int>(theta.size()) < kernel->GetParametersNumbers())
        throw std::domain_error("initial theta needs " + std::to_string(kernel->GetParametersNumbers()) + " values");
    egs_problem_t p = aData->GetProblem(ExaGeoStatHardware::GetContext(), kernel->GetDeviceKernelId(),
                                        aConfigurations.GetDistanceMetric(), aConfigurations.GetDenseTileSize());
    std::vector<double> &obs = aData->GetMeasurements();
    obs.assign(static_cast<size_t>(n) * kernel->GetVariablesNumber(), 0.0);
    CheckStatus(egs_generate_observations(p, theta.data(), aConfigurations.GetSeed(), obs.data()),
                "GenerateObservationsVector");
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
    aConfigurations.InitializeAllTheta();
    const int np = kernel->GetParametersNumbers();
    const size_t nz = static_cast<size_t>(aData->GetLocations()->GetSize()) * kernel->GetVariablesNumber();
    std::vector<double> &obs = aData->GetMeasurements();
    if (apMeasurementsMatrix) obs = ToDouble(apMeasurementsMatrix, nz);
    if (obs.size() != nz) throw std::runtime_error("ExaGeoStatDataModeling: no measurements");
    egs_problem_t p = aData->GetProblem(ExaGeoStatHardware::GetContext(), kernel->GetDeviceKernelId(),
                                        aConfigurations.GetDistanceMetric(), aConfigurations.GetDenseTileSize());
    CheckStatus(egs_problem_set_observations(p, obs.data(), (int64_t) nz), "egs_problem_set_observations");
    ConfigurePrecision(p, aConfigurations);

    std::vector<double> lb = aConfigurations.GetLowerBounds(), ub = aConfigurations.GetUpperBounds();
    std::vector<double> x = aConfigurations.GetStartingTheta();
    std::vector<int> free_idx;
    for (int i = 0; i < np; ++i) {
        if (ub[i] > lb[i]) free_idx.push_back(i);
        else if (ub[i] < lb[i]) throw std::domain_error("upper bound below lower bound for parameter " + std::to_string(i));
        x[i] = std::min(ub[i], std::max(lb[i], x[i]));
    }

    // One evaluation of ExaGeoStatMLETile (ChameleonImplementation.cpp:35-327):
    // llh, its time and flops, and the reference's per-evaluation summary keys.
    struct State {
        egs_problem_t p;
        std::vector<double> theta;
        const std::vector<int> *free_idx;
        ExaGeoStatData<T> *data;
        double n, total_time = 0.0, total_flops = 0.0, best = -INFINITY;
        std::vector<double> best_theta;
        int status = EGS_SUCCESS;
        std::string error;
        int count = 0;
    } st{p, x, &free_idx, aData.get(), static_cast<double>(nz)};
    auto objective = [](const double *v, void *user) -> double {
        auto *s = static_cast<State *>(user);
        for (size_t j = 0; j < s->free_idx->size(); ++j) s->theta[(*s->free_idx)[j]] = v[j];
        double llh = 0.0;
        const double t0 = Seconds();
        const int code = egs_loglik(s->p, s->theta.data(), &llh, nullptr, nullptr);
        const double dt = Seconds() - t0;
        if (code != EGS_SUCCESS) {
            if (s->status == EGS_SUCCESS) {
                s->status = code;
                s->error = egs_last_error();
            }
            return NAN;
        }
        ++s->count;
        s->total_time += dt;
        s->total_flops += (s->n * s->n * s->n / 3.0 + 2.0 * s->n * s->n) / 1e9;
        s->data->SetMleIterations(s->data->GetMleIterations() + 1);
        auto *res = results::Results::GetInstance();
        res->SetMLEIterations(s->data->GetMleIterations());
        res->SetMaximumTheta(s->theta);
        res->SetLogLikValue(llh);
        res->SetTotalModelingExecutionTime(s->total_time);
        res->SetTotalModelingFlops(s->total_flops);
        if (Verbose()) {
            std::printf("\t%d- Model Parameters (", s->data->GetMleIterations());
            for (size_t i = 0; i < s->theta.size(); ++i)
                std::printf("%.8f%s", s->theta[i], i + 1 < s->theta.size() ? ", " : "");
            std::printf(")----> LogLi: %.18f\n", llh);
        }
        if (llh > s->best) {
            s->best = llh;
            s->best_theta = s->theta;
        }
        return -llh;
    };
    if (free_idx.empty()) {
        objective(nullptr, &st);
    } else {
        const int nf = static_cast<int>(free_idx.size());
        std::vector<double> xf(nf), lf(nf), uf(nf);
        for (int j = 0; j < nf; ++j) {
            xf[j] = x[free_idx[j]];
            lf[j] = lb[free_idx[j]];
            uf[j] = ub[free_idx[j]];
        }
        double fmin = 0.0;
        int nev = 0, reason = 0;
        egs_bobyqa(nf, xf.data(), lf.data(), uf.data(), aConfigurations.GetTolerance(),
                   aConfigurations.GetMaxMleIterations(), 0.0, objective, &st, &fmin, &nev, &reason);
    }
    if (st.status != EGS_SUCCESS) throw std::runtime_error("ExaGeoStatMLETile failed: " + st.error);
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

    // ExaGeoStatMLEPredictTile (LinearAlgebraMethods.cpp:451-605) on the GPU.
    const double t0 = Seconds();
    auto ox = ToDouble(obs_loc.GetLocationX(), n_obs), oy = ToDouble(obs_loc.GetLocationY(), n_obs);
    auto oz = ToDouble(obs_loc.GetLocationZ(), dim == common::Dimension2D ? 0 : n_obs);
    auto mx = ToDouble(miss.GetLocationX(), z_miss), my = ToDouble(miss.GetLocationY(), z_miss);
    auto mz = ToDouble(miss.GetLocationZ(), dim == common::Dimension2D ? 0 : z_miss);
    egs_problem_t op = nullptr;
    CheckStatus(egs_problem_create(ExaGeoStatHardware::GetContext(), kernel->GetDeviceKernelId(), n_obs, ox.data(),
                                   oy.data(), oz.empty() ? nullptr : oz.data(), aConfigurations.GetDistanceMetric(),
                                   aConfigurations.GetDenseTileSize(), &op),
                "egs_problem_create");
    std::unique_ptr<egs_problem_s, int (*)(egs_problem_t)> guard(op, egs_problem_destroy);
    CheckStatus(egs_problem_set_observations(op, z_obs.data(), (int64_t) z_obs.size()), "egs_problem_set_observations");
    ConfigurePrecision(op, aConfigurations);
    std::vector<double> zmiss(static_cast<size_t>(z_miss) * p);
    double mspe = -1.0;
    std::vector<double> theta(aConfigurations.GetEstimatedTheta().begin(),
                              aConfigurations.GetEstimatedTheta().begin() + np);
    CheckStatus(egs_predict(op, theta.data(), z_miss, mx.data(), my.data(), mz.empty() ? nullptr : mz.data(),
                            zmiss.data(), z_actual.empty() ? nullptr : z_actual.data(), &mspe),
                "ExaGeoStatMLEPredictTile");
    const double dt = Seconds() - t0;
    auto *res = results::Results::GetInstance();
    res->SetMSPEExecutionTime(dt);
    const double no = static_cast<double>(n_obs) * p;
    res->SetMSPEFlops((no * no * no / 3.0 + 2.0 * no * z_miss * p) / 1e9 / std::max(dt, 1e-12));
    if (!z_actual.empty()) {
        res->SetMSPEError(mspe);
        if (Verbose()) std::printf("\t\t- Prediction value: %f\n", mspe);
    }
    res->SetPredictedMissedValues(zmiss);
}

template class ExaGeoStat<float>;
template class ExaGeoStat<double>;

}  // namespace api
}  // namespace exageostat
