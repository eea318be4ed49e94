// ExaGeoStat.hpp -- C++ host API of the MI355X likelihood path.
//
// Mirrors the reference's public C++ surface for the maximum-likelihood path
// so that a caller of ExaGeoStatCPP compiles against it unchanged:
//   exageostat::api::ExaGeoStat<T>          (inst/include/api/ExaGeoStat.hpp:29-80)
//   exageostat::configurations::Configurations (inst/include/configurations/Configurations.hpp:115-313)
//   exageostat::dataunits::Locations<T>, ExaGeoStatData<T> (inst/include/data-units/)
//   exageostat::kernels::Kernel<T> + PluginRegistry (inst/include/kernels/Kernel.hpp:45-140,
//                                                   inst/include/common/PluginRegistry.hpp:30-80)
//   ExaGeoStatHardware                      (inst/include/hardware/ExaGeoStatHardware.hpp)
//   exageostat::results::Results            (inst/include/results/Results.hpp:25-360)
// Every numerical step runs on the GPU through the C ABI of
// include/exageostat_amd.h; this layer owns no device code.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/exageostat_amd.h"

namespace exageostat {

namespace common {
// Values as in inst/include/common/Definitions.hpp.
enum Verbose { QUIET_MODE = 0, STANDARD_MODE = 1, DETAILED_MODE = 2 };
enum Dimension { Dimension2D = 0, Dimension3D = 1, DimensionST = 2 };
enum DistanceMetric { EUCLIDEAN_DISTANCE = 0, GREAT_CIRCLE_DISTANCE = 1 };
enum Computation { EXACT_DENSE = 0, DIAGONAL_APPROX = 1, TILE_LOW_RANK = 2 };
enum Precision { SINGLE = 0, DOUBLE = 1, MIXED = 2 };
enum UpperLower : int { EXAGEOSTAT_UPPER = 121, EXAGEOSTAT_LOWER = 122, EXAGEOSTAT_UPPER_LOWER = 123 };
enum Side { EXAGEOSTAT_LEFT = 141, EXAGEOSTAT_RIGHT = 142 };
enum Trans { EXAGEOSTAT_NO_TRANS = 111, EXAGEOSTAT_TRANS = 112, EXAGEOSTAT_CONJ_TRANS = 113 };
enum Diag { EXAGEOSTAT_NON_UNIT = 131, EXAGEOSTAT_UNIT = 132 };
}  // namespace common

// Throws std::runtime_error with the library's message when status != 0
// (the reference throws after a failed CHAMELEON_* call, ChameleonDense.cpp:30-32).
void CheckStatus(int aStatus, const char *aWhere);

namespace configurations {

// The configuration keys on the likelihood path, with the reference's
// defaults (configurations/config.json and Configurations.cpp:32-66).
class Configurations {
public:
    Configurations();

    void SetProblemSize(int aProblemSize) { mProblemSize = aProblemSize; }
    int GetProblemSize() const { return mProblemSize; }
    void SetKernelName(const std::string &aKernel) { mKernelName = aKernel; }
    const std::string &GetKernelName() const { return mKernelName; }
    void SetTimeSlot(int aTimeSlot) { mTimeSlot = aTimeSlot; }
    int GetTimeSlot() const { return mTimeSlot; }
    void SetComputation(common::Computation aComputation) { mComputation = aComputation; }
    common::Computation GetComputation() const { return mComputation; }
    void SetPrecision(common::Precision aPrecision) { mPrecision = aPrecision; }
    common::Precision GetPrecision() const { return mPrecision; }
    void SetCoresNumber(int aCores) { mCores = aCores; }
    int GetCoresNumber() const { return mCores; }
    void SetGPUsNumbers(int aGpus) { mGpus = aGpus; }
    int GetGPUsNumbers() const { return mGpus; }
    void SetDenseTileSize(int aTileSize) { mDts = aTileSize; }
    int GetDenseTileSize() const { return mDts; }
    void SetBand(int aBand) { mBand = aBand; }
    int GetBand() const { return mBand; }
    // Tile-distance bands of the mixed plan: FP64 up to DenseBandDP, FP32 up
    // to BandDenseSP, 16-bit products beyond (row-scaled split FP16, one FP16
    // product for blocks with small panel rows; keys "banddensedp"/"banddensesp").
    // Defaults 4 and 4: no FP32 band, which measured both faster and at least as
    // accurate as an FP32 band at n = 100k.  FP64 storage up to band 4 keeps
    // the 40k fit's estimate within 1e-4 of the FP64 one (band 1: 1.7e-3);
    // only the diagonal tiles take FP64 products (two bands of tiles for the
    // bivariate kernel).
    void SetDenseBandDP(int aBand) { mBandDP = aBand; }
    int GetDenseBandDP() const { return mBandDP; }
    void SetBandDenseSP(int aBand) { mBandSP = aBand; }
    int GetBandDenseSP() const { return mBandSP; }
    void SetSeed(int aSeed) { mSeed = aSeed; }
    int GetSeed() const { return mSeed; }
    void SetInitialTheta(const std::vector<double> &apTheta) { mInitialTheta = apTheta; }
    std::vector<double> &GetInitialTheta() { return mInitialTheta; }
    void SetLowerBounds(const std::vector<double> &apTheta) { mLowerBounds = apTheta; }
    std::vector<double> &GetLowerBounds() { return mLowerBounds; }
    void SetUpperBounds(const std::vector<double> &apTheta) { mUpperBounds = apTheta; }
    std::vector<double> &GetUpperBounds() { return mUpperBounds; }
    void SetEstimatedTheta(const std::vector<double> &apTheta) { mEstimatedTheta = apTheta; }
    std::vector<double> &GetEstimatedTheta() { return mEstimatedTheta; }
    void SetStartingTheta(const std::vector<double> &apTheta) { mStartingTheta = apTheta; }
    std::vector<double> &GetStartingTheta() { return mStartingTheta; }
    void SetDimension(common::Dimension aDimension) { mDimension = aDimension; }
    common::Dimension GetDimension() const { return mDimension; }
    void SetIsSynthetic(bool aIsSynthetic) { mIsSynthetic = aIsSynthetic; }
    bool GetIsSynthetic() const { return mIsSynthetic; }
    void SetDistanceMetric(common::DistanceMetric aMetric) { mMetric = aMetric; }
    common::DistanceMetric GetDistanceMetric() const { return mMetric; }
    void SetMaxMleIterations(int aIterations) { mMaxMleIterations = aIterations; }
    int GetMaxMleIterations() const { return mMaxMleIterations; }
    // The tolerance is given as an exponent k and stored as 10^-k (Configurations.cpp:380-382).
    void SetTolerance(double aTolerance);
    double GetTolerance() const { return mTolerance; }
    void SetUnknownObservationsNb(int aZMiss) { mZMiss = aZMiss; }
    int GetUnknownObservationsNb() const { return mZMiss; }
    void SetObservationNumber(int aNumber) { mObservationNumber = aNumber; }
    int GetObservationNumber() const { return mObservationNumber; }
    void SetIsMSPE(bool a) { mIsMSPE = a; }
    bool GetIsMSPE() const { return mIsMSPE; }
    void SetIsIDW(bool a) { mIsIDW = a; }
    bool GetIsIDW() const { return mIsIDW; }
    void SetIsMLOEMMOM(bool a) { mIsMLOEMMOM = a; }
    bool GetIsMLOEMMOM() const { return mIsMLOEMMOM; }
    void SetIsFisher(bool a) { mIsFisher = a; }
    bool GetIsFisher() const { return mIsFisher; }
    void SetIsNonGaussian(bool a) { mIsNonGaussian = a; }
    bool GetIsNonGaussian() const { return mIsNonGaussian; }
    // Real data: "x,y[,z],z_1[,z_2]" lines read by ExaGeoStatLoadData when the data is not synthetic.
    void SetDataPath(const std::string &aPath) { mDataPath = aPath; }
    const std::string &GetDataPath() const { return mDataPath; }
    // Synthetic data is also written to <LoggerPath>/synthetic_ds/SYN_<n>_<i> when Logger is on.
    void SetLogger(bool a) { mLogger = a; }
    bool GetLogger() const { return mLogger; }
    void SetLoggerPath(const std::string &aPath) { mLoggerPath = aPath; }
    const std::string &GetLoggerPath() const { return mLoggerPath; }

    static void SetVerbosity(common::Verbose aVerbose) { mVerbosity = aVerbose; }
    static common::Verbose GetVerbosity() { return mVerbosity; }

    // Configurations::InitializeAllTheta (Configurations.cpp:182-213): pads
    // every theta vector to the kernel's parameter count (-1 when empty, 0
    // otherwise), starts at the lower bounds, and pins every parameter whose
    // estimated value is set (!= -1) by collapsing its bounds onto it.
    void InitializeAllTheta();
    // n - zmiss (Configurations.cpp:364-366).
    int CalculateZObsNumber() const { return mProblemSize - mZMiss; }

private:
    static void InitTheta(std::vector<double> &aTheta, int aSize);

    int mProblemSize = 16;
    std::string mKernelName = "UnivariateMaternStationary";
    int mTimeSlot = 1;
    common::Computation mComputation = common::EXACT_DENSE;
    common::Precision mPrecision = common::DOUBLE;
    int mCores = 1, mGpus = 0, mDts = 8, mBand = 0, mBandDP = 4, mBandSP = 4;
    int mSeed = 0;
    std::vector<double> mInitialTheta, mLowerBounds, mUpperBounds, mEstimatedTheta, mStartingTheta;
    common::Dimension mDimension = common::Dimension2D;
    bool mIsSynthetic = true;
    common::DistanceMetric mMetric = common::EUCLIDEAN_DISTANCE;
    int mMaxMleIterations = 1;
    double mTolerance = 0.1;
    int mZMiss = 0, mObservationNumber = 0;
    bool mIsMSPE = false, mIsIDW = false, mIsMLOEMMOM = false, mIsFisher = false, mIsNonGaussian = false;
    std::string mDataPath, mLoggerPath;
    bool mLogger = false;
    bool mIsThetaInit = false;
    static common::Verbose mVerbosity;
};

}  // namespace configurations

namespace dataunits {

// Structure-of-arrays locations (Locations.cpp:94-117).
template<typename T>
class Locations {
public:
    Locations(int aSize, common::Dimension aDimension)
        : mSize(aSize), mDimension(aDimension), mX(aSize), mY(aSize),
          mZ(aDimension == common::Dimension2D ? 0 : aSize) {}
    void SetLocationX(T &aLocationX, const int &aSize) { mX.assign(&aLocationX, &aLocationX + aSize); }
    void SetLocationY(T &aLocationY, const int &aSize) { mY.assign(&aLocationY, &aLocationY + aSize); }
    void SetLocationZ(T &aLocationZ, const int &aSize) { mZ.assign(&aLocationZ, &aLocationZ + aSize); }
    T *GetLocationX() { return mX.data(); }
    T *GetLocationY() { return mY.data(); }
    T *GetLocationZ() { return mZ.empty() ? nullptr : mZ.data(); }
    void SetSize(const int &aSize) { mSize = aSize; }
    int GetSize() const { return mSize; }
    void SetDimension(const common::Dimension &aDimension) { mDimension = aDimension; }
    common::Dimension GetDimension() const { return mDimension; }

private:
    int mSize;
    common::Dimension mDimension;
    std::vector<T> mX, mY, mZ;
};

}  // namespace dataunits
}  // namespace exageostat

// Locations, observations and the device-resident problem built from them,
// in the global namespace as in the reference (ExaGeoStatData.hpp:25-110,
// ExaGeoStatData.cpp:21-38 with DescriptorData's role taken by the C ABI's
// egs_problem_t: tiles, factor and observations stay in HBM between calls).
template<typename T>
class ExaGeoStatData {
public:
    ExaGeoStatData(int aSize, exageostat::common::Dimension aDimension)
        : mLocations(std::make_unique<exageostat::dataunits::Locations<T>>(aSize, aDimension)) {}
    ~ExaGeoStatData() { ReleaseProblem(); }
    ExaGeoStatData(const ExaGeoStatData &) = delete;
    ExaGeoStatData &operator=(const ExaGeoStatData &) = delete;

    exageostat::dataunits::Locations<T> *GetLocations() { return mLocations.get(); }
    void SetLocations(exageostat::dataunits::Locations<T> &aLocation) {
        mLocations = std::make_unique<exageostat::dataunits::Locations<T>>(aLocation);
        ReleaseProblem();
    }
    void SetMleIterations(const int &aMleIterations) { mMleIterations = aMleIterations; }
    int GetMleIterations() const { return mMleIterations; }

    // Observations (length n * variables); empty until generated or set.
    std::vector<double> &GetMeasurements() { return mZ; }
    // Device problem for (kernel, metric, dts); rebuilt when any of them changes.
    egs_problem_t GetProblem(egs_context_t aContext, int aKernelId, int aMetric, int aDts);
    void ReleaseProblem();
    // Sets the problem's precision plan (mode, FP64 band, FP32 band) unless it
    // is already in effect; a plan change re-lays out the tiles on the device.
    void SetProblemPlan(int aMode, int aBandDP, int aBandSP);

private:
    std::unique_ptr<exageostat::dataunits::Locations<T>> mLocations;
    int mMleIterations = 0;
    std::vector<double> mZ;
    egs_problem_t mProblem = nullptr;
    int mKernelId = -1, mMetric = -1, mDts = -1;
    int mPlan[3] = {-1, -1, -1};
};


namespace exageostat {
namespace kernels {

struct KernelsConfigurations {
    // name -> number of parameters, filled when a kernel registers.
    static std::unordered_map<std::string, int> &GetParametersNumberKernelMap();
};

// Covariance plugin interface (Kernel.hpp:45-140).  The concrete kernels
// evaluate their tiles on the GPU (egs_covariance_tile).
template<typename T>
class Kernel {
public:
    virtual ~Kernel() = default;
    virtual void GenerateCovarianceMatrix(T *apMatrixA, const int &aRowsNumber, const int &aColumnsNumber,
                                          const int &aRowOffset, const int &aColumnOffset,
                                          dataunits::Locations<T> &aLocation1, dataunits::Locations<T> &aLocation2,
                                          dataunits::Locations<T> &aLocation3, T *apLocalTheta,
                                          const int &aDistanceMetric) = 0;
    int GetVariablesNumber() const { return mVariablesNumber; }
    void SetPValue(int aTimeSlot) { mVariablesNumber = mP * aTimeSlot; }
    int GetParametersNumbers() const { return mParametersNumber; }
    // Kernel id of the C ABI (EGS_KERNEL_*).
    virtual int GetDeviceKernelId() const = 0;

protected:
    int mP = 1;
    int mVariablesNumber = 1;
    int mParametersNumber = 3;
};

// UnivariateMaternStationary.cpp:39-66: theta = (sigma^2, range, nu).
template<typename T>
class UnivariateMaternStationary : public Kernel<T> {
public:
    UnivariateMaternStationary() {
        this->mP = 1;
        this->mVariablesNumber = 1;
        this->mParametersNumber = 3;
    }
    void GenerateCovarianceMatrix(T *apMatrixA, const int &aRowsNumber, const int &aColumnsNumber,
                                  const int &aRowOffset, const int &aColumnOffset,
                                  dataunits::Locations<T> &aLocation1, dataunits::Locations<T> &aLocation2,
                                  dataunits::Locations<T> &aLocation3, T *apLocalTheta,
                                  const int &aDistanceMetric) override;
    int GetDeviceKernelId() const override { return EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY; }
};

// BivariateMaternParsimonious.cpp:40-108: theta = (sigma1^2, sigma2^2, range, nu1, nu2, beta).
template<typename T>
class BivariateMaternParsimonious : public Kernel<T> {
public:
    BivariateMaternParsimonious() {
        this->mP = 2;
        this->mVariablesNumber = 2;
        this->mParametersNumber = 6;
    }
    void GenerateCovarianceMatrix(T *apMatrixA, const int &aRowsNumber, const int &aColumnsNumber,
                                  const int &aRowOffset, const int &aColumnOffset,
                                  dataunits::Locations<T> &aLocation1, dataunits::Locations<T> &aLocation2,
                                  dataunits::Locations<T> &aLocation3, T *apLocalTheta,
                                  const int &aDistanceMetric) override;
    int GetDeviceKernelId() const override { return EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS; }
};

}  // namespace kernels

namespace plugins {

// Self-registering factory (PluginRegistry.hpp:30-80): Create returns an owning
// raw pointer, or nullptr for an unknown name.
template<typename T>
class PluginRegistry {
public:
    typedef std::function<T *()> FactoryFunction;
    typedef std::unordered_map<std::string, FactoryFunction> FactoryMap;

    static bool Add(const std::string &name, FactoryFunction fac) {
        auto &map = GetFactoryMap();
        if (map.find(name) != map.end()) return false;
        map[name] = std::move(fac);
        return true;
    }
    static T *Create(const std::string &aName, const int &aTimeSlot) {
        auto &map = GetFactoryMap();
        auto it = map.find(aName);
        if (it == map.end()) return nullptr;
        T *object = it->second();
        object->SetPValue(aTimeSlot);
        return object;
    }

private:
    static FactoryMap &GetFactoryMap() {
        static FactoryMap mSelfRegisteringMap;
        return mSelfRegisteringMap;
    }
};

}  // namespace plugins

namespace dataLoader {
namespace csv {

// CSV data files (src/data-loader/concrete/CSVLoader.cpp:36-234): one location
// per line, "x,y,z_1..z_P" in 2D and "x,y,zloc,z_1..z_P" in 3D.
template<typename T>
class CSVLoader {
public:
    static CSVLoader<T> *GetInstance();
    // Reads the file at aConfigurations.GetDataPath() and sets the problem size
    // to lines * P / time slot; throws std::runtime_error with the reference's
    // messages for a missing path, an unreadable file or inconsistent columns.
    void ReadData(configurations::Configurations &aConfigurations, std::vector<T> &aMeasurementsMatrix,
                  std::vector<T> &aXLocations, std::vector<T> &aYLocations, std::vector<T> &aZLocations,
                  const int &aP);
    // Writes <aLoggerPath>/synthetic_ds/SYN_<n/P>_<i> with 15 significant digits,
    // i the first index whose log_<n/P>_<i> file does not exist.
    void WriteData(const T &aMatrixPointer, const int &aProblemSize, const int &aP, std::string &aLoggerPath,
                   dataunits::Locations<T> &aLocations);
};

}  // namespace csv
}  // namespace dataLoader

namespace results {

// Run summary (Results.cpp:20-200): the keys of the likelihood path, printed
// as "#<key>: <value>" lines by PrintEndSummary.
class Results {
public:
    static Results *GetInstance();
    void SetIsSynthetic(bool aIsSynthetic, const std::string &aKey = "");
    void SetGeneratedLocationsNumber(int aNumLocations, const std::string &aKey = "");
    void SetTotalDataGenerationExecutionTime(double aTime, const std::string &aKey = "");
    void SetTotalDataGenerationFlops(double aFlops, const std::string &aKey = "");
    void SetLogLikValue(double aLogLikValue, const std::string &aKey = "");
    void SetMLEIterations(int aIterationsNumber, const std::string &aKey = "");
    void SetMaximumTheta(const std::vector<double> &aMaximumTheta, const std::string &aKey = "");
    void SetTotalModelingExecutionTime(double aTime, const std::string &aKey = "");
    void SetTotalModelingFlops(double aFlops, const std::string &aKey = "");
    void SetZMiss(int aZMiss, const std::string &aKey = "");
    void SetMSPEError(double aMSPEError, const std::string &aKey = "");
    void SetMSPEExecutionTime(double aTime, const std::string &aKey = "");
    void SetMSPEFlops(double aFlops, const std::string &aKey = "");
    void SetPredictedMissedValues(std::vector<double> aPredictedValues, const std::string &aKey = "");

    double GetLogLikValue() const { return mLogLikValue; }
    int GetMLEIterations() const { return mMLEIterations; }
    std::vector<double> GetMaximumTheta() const { return mMaximumTheta; }
    double GetTotalModelingExecutionTime() const { return mTotalModelingTime; }
    double GetTotalModelingFlops() const { return mTotalModelingFlops; }
    double GetAverageModelingExecutionTime() const;
    double GetAverageModelingFlops() const;
    double GetMSPEError() const { return mMSPEError; }
    std::vector<double> GetPredictedMissedValues() const { return mPredicted; }
    const std::map<std::string, std::string> &GetSummary() const { return mSummaryDictionary; }
    void PrintEndSummary() const;

private:
    void UpdateDictionary(const std::string &key, const std::string &value) { mSummaryDictionary[key] = value; }
    double mLogLikValue = 0.0, mTotalModelingTime = 0.0, mTotalModelingFlops = 0.0, mMSPEError = 0.0;
    int mMLEIterations = 0;
    std::vector<double> mMaximumTheta, mPredicted;
    std::map<std::string, std::string> mSummaryDictionary;
};

}  // namespace results

}  // namespace exageostat

namespace exageostat {
namespace dataunits {

// A vector descriptor of the dense solver: caller memory and its length (the
// reference's Z descriptor aliases user memory, LinearAlgebraMethods.cpp:75-77).
// Matrix descriptors are the device-resident problem (an egs_problem_t holding
// the lower tiles of Sigma, then of L).
template<typename T>
struct VectorDescriptor {
    T *data = nullptr;
    int64_t size = 0;
};

}  // namespace dataunits

namespace linearAlgebra {

// The solver interface (LinearAlgebraMethods.hpp:36-330): the reference's pure
// virtuals with their signatures (:140-142, :153, :161, :170, :185-186,
// :203-205), plus the three base-class operations that the reference
// implements with direct CHAMELEON_* calls (GenerateObservationsVector,
// ExaGeoStatPosvTile, ExaGeoStatMLEPredictTile; LinearAlgebraMethods.cpp:333-605,
// 1421-1428), made virtual here so that the concrete solver owns them.
// Descriptors (void *) are egs_problem_t for matrices and
// dataunits::VectorDescriptor<T> * for vectors; a sequence is an
// egs_context_t.  Failures throw std::runtime_error.
template<typename T>
class LinearAlgebraMethods {
public:
    virtual ~LinearAlgebraMethods() = default;

    virtual T ExaGeoStatMLETile(std::unique_ptr<ExaGeoStatData<T>> &aData,
                                configurations::Configurations &aConfigurations, const double *apTheta,
                                T *apMeasurementsMatrix, const kernels::Kernel<T> &aKernel) = 0;
    virtual void ExaGeoStatLapackCopyTile(const common::UpperLower &aUpperLower, void *apA, void *apB) = 0;
    virtual void ExaGeoStatSequenceWait(void *apSequence) = 0;
    virtual void ExaGeoStatCreateSequence(void *apSequence) = 0;
    virtual void ExaGeoStatPotrfTile(const common::UpperLower &aUpperLower, void *apA, int aBand, void *apCD,
                                     void *apCrk, const int &aMaxRank, const int &aAcc) = 0;
    virtual void ExaGeoStatTrsmTile(const common::Side &aSide, const common::UpperLower &aUpperLower,
                                    const common::Trans &aTrans, const common::Diag &aDiag, const T &aAlpha,
                                    void *apA, void *apCD, void *apCrk, void *apZ, const int &aMaxRank) = 0;

    virtual void GenerateObservationsVector(configurations::Configurations &aConfigurations,
                                            std::unique_ptr<ExaGeoStatData<T>> &aData,
                                            dataunits::Locations<T> *apLocation1, dataunits::Locations<T> *apLocation2,
                                            dataunits::Locations<T> *apLocation3, const int &aDistanceMetric,
                                            const kernels::Kernel<T> &aKernel) = 0;
    virtual void ExaGeoStatPosvTile(const common::UpperLower &aUpperLower, void *apA, void *apB) = 0;
    // aZMissNumber / aZObsNumber count locations; the vectors hold P values per
    // location.  Returns new T[3] = {mspe, mspe1, mspe2} (the caller deletes
    // it), as LinearAlgebraMethods.cpp:596-600 does.
    virtual T *ExaGeoStatMLEPredictTile(std::unique_ptr<ExaGeoStatData<T>> &aData, T *apTheta,
                                        const int &aZMissNumber, const int &aZObsNumber, T *apZObs, T *apZActual,
                                        T *apZMiss, configurations::Configurations &aConfiguration,
                                        dataunits::Locations<T> &aMissLocations,
                                        dataunits::Locations<T> &aObsLocations,
                                        const kernels::Kernel<T> &aKernel) = 0;
};

namespace dense {

// The EXACT_DENSE solver on the MI355X: the role of ChameleonDense<T> with
// ChameleonImplementation<T> (ChameleonDense.cpp:25-37,
// ChameleonImplementation.cpp:35-365), every step on the GPU through the C ABI.
template<typename T>
class HipDense : public LinearAlgebraMethods<T> {
public:
    T ExaGeoStatMLETile(std::unique_ptr<ExaGeoStatData<T>> &aData, configurations::Configurations &aConfigurations,
                        const double *apTheta, T *apMeasurementsMatrix, const kernels::Kernel<T> &aKernel) override;
    void ExaGeoStatLapackCopyTile(const common::UpperLower &aUpperLower, void *apA, void *apB) override;
    void ExaGeoStatSequenceWait(void *apSequence) override;
    void ExaGeoStatCreateSequence(void *apSequence) override;
    void ExaGeoStatPotrfTile(const common::UpperLower &aUpperLower, void *apA, int aBand, void *apCD, void *apCrk,
                             const int &aMaxRank, const int &aAcc) override;
    void ExaGeoStatTrsmTile(const common::Side &aSide, const common::UpperLower &aUpperLower,
                            const common::Trans &aTrans, const common::Diag &aDiag, const T &aAlpha, void *apA,
                            void *apCD, void *apCrk, void *apZ, const int &aMaxRank) override;
    void GenerateObservationsVector(configurations::Configurations &aConfigurations,
                                    std::unique_ptr<ExaGeoStatData<T>> &aData, dataunits::Locations<T> *apLocation1,
                                    dataunits::Locations<T> *apLocation2, dataunits::Locations<T> *apLocation3,
                                    const int &aDistanceMetric, const kernels::Kernel<T> &aKernel) override;
    void ExaGeoStatPosvTile(const common::UpperLower &aUpperLower, void *apA, void *apB) override;
    T *ExaGeoStatMLEPredictTile(std::unique_ptr<ExaGeoStatData<T>> &aData, T *apTheta, const int &aZMissNumber,
                                const int &aZObsNumber, T *apZObs, T *apZActual, T *apZMiss,
                                configurations::Configurations &aConfiguration,
                                dataunits::Locations<T> &aMissLocations, dataunits::Locations<T> &aObsLocations,
                                const kernels::Kernel<T> &aKernel) override;
};

}  // namespace dense

// LinearAlgebraFactory.cpp:31-54: EXACT_DENSE gives the MI355X dense solver;
// the tile-low-rank and diagonal-super-tile solvers are not on this path.
template<typename T>
class LinearAlgebraFactory {
public:
    static std::unique_ptr<LinearAlgebraMethods<T>> CreateLinearAlgebraSolver(common::Computation aComputation);
};

}  // namespace linearAlgebra
}  // namespace exageostat

// Runtime bring-up (ExaGeoStatHardware.hpp): one GPU context shared by the
// API calls while the object lives.  The reference's core count drives
// StarPU workers; here the host side is single-threaded and the GPU count may
// be 0 or 1 (the path runs on GPU 0; more than one GPU is rejected).
class ExaGeoStatHardware {
public:
    ExaGeoStatHardware(const exageostat::common::Computation &aComputation, const int &aCoreNumber,
                       const int &aGpuNumber);
    ~ExaGeoStatHardware();
    ExaGeoStatHardware(const ExaGeoStatHardware &) = delete;
    ExaGeoStatHardware &operator=(const ExaGeoStatHardware &) = delete;
    ExaGeoStatHardware(ExaGeoStatHardware &&aOther) noexcept;
    // The live context (throws when no ExaGeoStatHardware exists).
    static egs_context_t GetContext();

private:
    bool mOwner = false;
};

namespace exageostat {
namespace api {

// The three entry points of ExaGeoStat.hpp:29-80.
template<typename T>
class ExaGeoStat {
public:
    // ExaGeoStat.cpp:73-88: srand(seed), generate locations (Morton-ordered,
    // 16-bit snapped), then z = L(initial theta) e with the dlarnv stream
    // (the solver's GenerateObservationsVector); real data is read from the
    // configured CSV file.
    static void ExaGeoStatLoadData(configurations::Configurations &aConfigurations,
                                   std::unique_ptr<ExaGeoStatData<T>> &aData);
    // Synthetic data generation: ExaGeoStatLoadData's synthetic branch,
    // whatever the configuration's data path says.
    static void ExaGeoStatGenerateData(configurations::Configurations &aConfigurations,
                                       std::unique_ptr<ExaGeoStatData<T>> &aData);
    // ExaGeoStat.cpp:91-107 + StarpuRuntimeSolver.cpp:28-79: BOBYQA on the
    // log-likelihood from the lower bounds, each evaluation the solver's
    // ExaGeoStatMLETile (the solver from LinearAlgebraFactory for the
    // configured computation); returns the best log-likelihood and stores the
    // best theta as the estimated theta.
    static T ExaGeoStatDataModeling(configurations::Configurations &aConfigurations,
                                    std::unique_ptr<ExaGeoStatData<T>> &aData, T *apMeasurementsMatrix = nullptr);
    // ExaGeoStat.cpp:111-125 + Prediction.cpp:35-210: kriging of the missing
    // observations and MSPE.  IDW, MLOE-MMOM and Fisher are not on this path
    // and raise std::runtime_error.
    static void ExaGeoStatPrediction(configurations::Configurations &aConfigurations,
                                     std::unique_ptr<ExaGeoStatData<T>> &aData, T *apMeasurementsMatrix = nullptr,
                                     dataunits::Locations<T> *apTrainLocations = nullptr,
                                     dataunits::Locations<T> *apTestLocations = nullptr);
};

}  // namespace api
}  // namespace exageostat
