// Tests of the C++ host API, written after the reference's own API tests
// (tests/cpp-tests/api/TestExaGeoStatApi.cpp, tests/cpp-tests/prediction/TestPrediction.cpp,
// tests/cpp-tests/Rcpp-adapters/TestAllRFunctions.cpp, tests/cpp-tests/configurations/).
//
//   test_exageostat_api cpu            host-only cases (no GPU needed)
//   test_exageostat_api gpu            the reference's API fixtures on the GPU
//   test_exageostat_api mle t0 t1 t2 llh
//        n = 1600 fit from the lower bounds; theta-hat and llh compared with
//        the optimum given on the command line (computed from the oracle)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <vector>

#include "../../exageostatcpp_amd/api/ExaGeoStat.hpp"

using namespace exageostat::common;
using namespace exageostat::dataunits;
using namespace exageostat::configurations;
using exageostat::dataunits::VectorDescriptor;

static int g_failures = 0;

#define REQUIRE(cond)                                                                   \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            std::printf("  requirement failed at line %d: %s\n", __LINE__, #cond);      \
            throw std::runtime_error("requirement failed");                             \
        }                                                                               \
    } while (0)

static void run(const char *name, const std::function<void()> &fn) {
    std::printf("CASE %s\n", name);
    std::fflush(stdout);
    try {
        fn();
        std::printf("PASS %s\n", name);
    } catch (const std::exception &e) {
        ++g_failures;
        std::printf("FAIL %s: %s\n", name, e.what());
    }
    std::fflush(stdout);
}

// Fixture data of TestExaGeoStatApi.cpp:111-133 (N = 16).
static double kZ16[16] = {-1.272336140360187606, -2.590699695867695773, 0.512142584178685967,
                          -0.163880452049749520, 0.313503633252489700, -1.474410682226017677,
                          0.161705025505231914, 0.623389205185149065, -1.341858445399783495,
                          -1.054282062428600009, -1.669383221392507943, 0.219170645803740793,
                          0.971213790000161170, 0.538973474182433021, -0.752828466476077041,
                          0.290822066007430102};
static double kX16[16] = {0.193041886015106440, 0.330556191348134576, 0.181612878614480805,
                          0.370473792629892440, 0.652140077821011688, 0.806332494087129037,
                          0.553322652018005678, 0.800961318379491916, 0.207324330510414295,
                          0.347951476310368490, 0.092042420080872822, 0.465445944914930965,
                          0.528267338063630132, 0.974792095826657490, 0.552452887769893985,
                          0.877592126344701295};
static double kY16[16] = {0.103883421072709245, 0.135790035858701447, 0.434683756771190977,
                          0.400778210116731537, 0.168459601739528508, 0.105195696955825133,
                          0.396398870832379624, 0.296757457846952011, 0.564507515068284116,
                          0.627679865720607300, 0.928648813611047563, 0.958236057068741931,
                          0.573571374074921758, 0.568657969024185528, 0.935835812924391552,
                          0.942824444953078489};

static void cpu_cases() {
    run("configurations defaults", [] {
        Configurations c;
        REQUIRE(c.GetProblemSize() == 16);
        REQUIRE(c.GetKernelName() == "UnivariateMaternStationary");
        REQUIRE(c.GetDenseTileSize() == 8);
        REQUIRE(c.GetMaxMleIterations() == 1);
        REQUIRE(c.GetPrecision() == DOUBLE);
        REQUIRE(c.GetDistanceMetric() == EUCLIDEAN_DISTANCE);
        c.SetTolerance(4);
        REQUIRE(std::fabs(c.GetTolerance() - 1e-4) < 1e-18);
    });
    run("InitializeAllTheta pads, starts at lb and pins estimated values", [] {
        Configurations c;
        c.SetLowerBounds({0.1, 0.2});
        c.SetUpperBounds({5, 5, 5});
        c.SetEstimatedTheta({-1, 0.3, -1});
        c.InitializeAllTheta();
        REQUIRE(c.GetInitialTheta().size() == 3 && c.GetInitialTheta()[0] == -1);  // empty -> -1
        REQUIRE(c.GetLowerBounds()[2] == 0.0 || c.GetLowerBounds()[2] == 0.3);
        REQUIRE(c.GetLowerBounds()[0] == 0.1);
        REQUIRE(c.GetLowerBounds()[1] == 0.3 && c.GetUpperBounds()[1] == 0.3);     // pinned
        REQUIRE(c.GetStartingTheta()[0] == 0.1 && c.GetStartingTheta()[1] == 0.3);
        REQUIRE(c.GetStartingTheta()[2] == 0.0);                                   // short lb padded with 0
    });
    run("plugin registry", [] {
        using exageostat::kernels::Kernel;
        using exageostat::plugins::PluginRegistry;
        std::unique_ptr<Kernel<double>> u(PluginRegistry<Kernel<double>>::Create("UnivariateMaternStationary", 1));
        REQUIRE(u && u->GetParametersNumbers() == 3 && u->GetVariablesNumber() == 1);
        std::unique_ptr<Kernel<double>> b(PluginRegistry<Kernel<double>>::Create("BivariateMaternParsimonious", 1));
        REQUIRE(b && b->GetParametersNumbers() == 6 && b->GetVariablesNumber() == 2);
        std::unique_ptr<Kernel<float>> bf(PluginRegistry<Kernel<float>>::Create("BivariateMaternParsimonious", 3));
        REQUIRE(bf && bf->GetVariablesNumber() == 6);                              // P * time slot
        REQUIRE(PluginRegistry<Kernel<double>>::Create("NoSuchKernel", 1) == nullptr);
        REQUIRE(!PluginRegistry<Kernel<double>>::Add("UnivariateMaternStationary", [] { return (Kernel<double> *) nullptr; }));
    });
    run("BOBYQA through the C ABI", [] {
        auto rosen = [](const double *v, void *) {
            return 100 * (v[1] - v[0] * v[0]) * (v[1] - v[0] * v[0]) + (1 - v[0]) * (1 - v[0]);
        };
        double x[2] = {-1.2, 1.0}, lb[2] = {-2, -2}, ub[2] = {2, 2}, fmin = 0;
        int nev = 0, reason = 0;
        REQUIRE(egs_bobyqa(2, x, lb, ub, 0.0, 3000, 1e-9, rosen, nullptr, &fmin, &nev, &reason) == 0);
        REQUIRE(std::fabs(x[0] - 1) < 1e-6 && std::fabs(x[1] - 1) < 1e-6 && fmin < 1e-12);
        double bad_lb[2] = {1, 1}, bad_ub[2] = {0, 2};
        REQUIRE(egs_bobyqa(2, x, bad_lb, bad_ub, 0.0, 10, 0, rosen, nullptr, &fmin, &nev, &reason) == EGS_ERR_INVALID_ARG);
    });
    run("CSV write then read round trip (CSVLoader.cpp:36-234)", [] {
        using exageostat::dataLoader::csv::CSVLoader;
        char tmpl[] = "/tmp/egs_csv_XXXXXX";
        REQUIRE(mkdtemp(tmpl) != nullptr);
        std::string dir = tmpl;
        Locations<double> l(16, Dimension2D);
        l.SetLocationX(*kX16, 16);
        l.SetLocationY(*kY16, 16);
        std::vector<double> z2(32);
        for (int i = 0; i < 32; ++i) z2[i] = kZ16[i % 16] * (i % 2 ? -1.0 : 1.0);
        std::string path = dir;
        CSVLoader<double>::GetInstance()->WriteData(*z2.data(), 32, 2, path, l);   // bivariate: P = 2
        REQUIRE(path == dir + "/synthetic_ds");
        Configurations c;
        c.SetDataPath(path + "/SYN_16_1");
        std::vector<double> z, x, y, zz;
        CSVLoader<double>::GetInstance()->ReadData(c, z, x, y, zz, 2);
        REQUIRE(c.GetProblemSize() == 32 && x.size() == 16 && zz.empty());
        for (int i = 0; i < 16; ++i) REQUIRE(std::fabs(x[i] - kX16[i]) < 1e-14 && std::fabs(y[i] - kY16[i]) < 1e-14);
        for (int i = 0; i < 32; ++i) REQUIRE(std::fabs(z[i] - z2[i]) < 1e-14);
        // two values per line cannot feed P = 3
        std::vector<double> a, b, d, e;
        Configurations c1;
        c1.SetDataPath(path + "/SYN_16_1");
        bool threw = false;
        try {
            CSVLoader<double>::GetInstance()->ReadData(c1, a, b, d, e, 3);
        } catch (const std::runtime_error &err) {
            threw = std::string(err.what()).find("isn't consistent") != std::string::npos;
        }
        REQUIRE(threw);
        Configurations c2;
        threw = false;
        try {
            CSVLoader<double>::GetInstance()->ReadData(c2, a, b, d, e, 1);
        } catch (const std::runtime_error &err) {
            threw = std::string(err.what()) == "Please enter the path to data file.";
        }
        REQUIRE(threw);
    });
    run("solver factory: EXACT_DENSE only (LinearAlgebraFactory.cpp:31-54)", [] {
        using exageostat::linearAlgebra::LinearAlgebraFactory;
        auto solver = LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        REQUIRE(solver != nullptr);
        REQUIRE(dynamic_cast<exageostat::linearAlgebra::dense::HipDense<double> *>(solver.get()) != nullptr);
        for (Computation c : {TILE_LOW_RANK, DIAGONAL_APPROX}) {
            bool threw = false;
            try {
                LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(c);
            } catch (const std::exception &) {
                threw = true;
            }
            REQUIRE(threw);
        }
    });
    run("solver: Z copy between vector descriptors (ChameleonImplementation.cpp:329-335)", [] {
        auto solver = exageostat::linearAlgebra::LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        std::vector<double> a(kZ16, kZ16 + 16), b(16, 0.0), c(3, 0.0);
        VectorDescriptor<double> da{a.data(), 16}, db{b.data(), 16}, dc{c.data(), 3};
        solver->ExaGeoStatLapackCopyTile(EXAGEOSTAT_UPPER_LOWER, &da, &db);
        REQUIRE(b == a);
        bool threw = false;
        try {
            solver->ExaGeoStatLapackCopyTile(EXAGEOSTAT_UPPER_LOWER, &da, &dc);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        REQUIRE(threw);
    });
    run("prediction without an estimated theta throws", [] {
        Configurations c;
        c.SetIsMSPE(true);
        c.SetUnknownObservationsNb(4);
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        bool threw = false;
        try {
            exageostat::api::ExaGeoStat<double>::ExaGeoStatPrediction(c, data, kZ16);
        } catch (const std::runtime_error &e) {
            threw = std::string(e.what()).find("estimated theta") != std::string::npos;
        }
        REQUIRE(threw);
    });
}

static void gpu_cases() {
    Configurations::SetVerbosity(QUIET_MODE);
    run("data generation - observations (TestExaGeoStatApi.cpp:23-63)", [] {
        Configurations c;
        const int N = 16;
        c.SetSeed(0);
        c.SetProblemSize(N);
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(9);
        c.SetComputation(EXACT_DENSE);
        c.SetLowerBounds({0.1, 0.1, 0.1});
        c.SetUpperBounds({5, 5, 5});
        c.SetInitialTheta({1, 0.1, 0.5});
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 4, 0);
        std::unique_ptr<ExaGeoStatData<double>> data;
        exageostat::api::ExaGeoStat<double>::ExaGeoStatLoadData(c, data);
        const double expected[] = {-1.272336, -2.590700, 0.512143, -0.163880, 0.313504, -1.474411, 0.161705, 0.623389,
                                   -1.341858, -1.054282, -1.669383, 0.219171, 0.971214, 0.538973, -0.752828, 0.290822};
        for (int i = 0; i < N; ++i) REQUIRE(std::fabs(data->GetMeasurements()[i] - expected[i]) < 1e-6);
        for (int i = 0; i < N; ++i) REQUIRE(std::fabs(data->GetLocations()->GetLocationX()[i] - kX16[i]) < 1e-15);
    });
    auto model_config = [](Configurations &c) {
        c.SetSeed(0);
        c.SetProblemSize(16);
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(8);
        c.SetComputation(EXACT_DENSE);
        c.SetMaxMleIterations(3);
        c.SetTolerance(4);
        c.SetLowerBounds({0.1, 0.1, 0.1});
        c.SetStartingTheta({0.1, 0.1, 0.1});
        c.SetUpperBounds({5, 5, 5});
        c.SetInitialTheta({1, 0.1, 0.5});
    };
    run("data modeling (TestExaGeoStatApi.cpp:103-139)", [&] {
        Configurations c;
        model_config(c);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 4, 0);
        auto data = std::make_unique<ExaGeoStatData<double>>(c.GetProblemSize(), c.GetDimension());
        data->GetLocations()->SetLocationX(*kX16, 16);
        data->GetLocations()->SetLocationY(*kY16, 16);
        double z[16];
        std::memcpy(z, kZ16, sizeof(z));
        double llh = exageostat::api::ExaGeoStat<double>::ExaGeoStatDataModeling(c, data, z);
        REQUIRE(std::fabs(llh - (-24.026000)) < 1e-6);
        REQUIRE(data->GetMleIterations() == 3);
        REQUIRE(exageostat::results::Results::GetInstance()->GetMLEIterations() == 3);
    });
    run("data generation and modeling (TestExaGeoStatApi.cpp:139-146)", [&] {
        Configurations c;
        model_config(c);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 4, 0);
        std::unique_ptr<ExaGeoStatData<double>> data;
        exageostat::api::ExaGeoStat<double>::ExaGeoStatLoadData(c, data);
        double llh = exageostat::api::ExaGeoStat<double>::ExaGeoStatDataModeling(c, data);
        REQUIRE(std::fabs(llh - (-24.026000)) < 1e-6);
        // the estimated theta is the best point of the three evaluations: lb + 1.225 e_1
        REQUIRE(std::fabs(c.GetEstimatedTheta()[0] - 1.325) < 1e-12 && c.GetEstimatedTheta()[1] == 0.1);
    });
    run("prediction MSPE (TestPrediction.cpp:77-95)", [] {
        Configurations c;
        c.SetUnknownObservationsNb(4);
        c.SetProblemSize(16);
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(8);
        c.SetInitialTheta({1, 0.1, 0.5});
        c.SetEstimatedTheta({0.9, 0.09, 0.4});
        c.SetIsMSPE(true);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        data->GetLocations()->SetLocationX(*kX16, 16);
        data->GetLocations()->SetLocationY(*kY16, 16);
        std::srand(1);  // the state of an unseeded rand(), as in the reference's test process
        exageostat::api::ExaGeoStat<double>::ExaGeoStatPrediction(c, data, kZ16);
        const double mspe = exageostat::results::Results::GetInstance()->GetMSPEError();
        REQUIRE(std::fabs(mspe - 0.552448) < 1e-5 * 0.552448);
        REQUIRE(exageostat::results::Results::GetInstance()->GetPredictedMissedValues().size() == 4);
    });
    run("prediction with train/test locations (TestAllRFunctions.cpp:64-84)", [] {
        Configurations c;
        c.SetProblemSize(16);
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(8);
        c.SetEstimatedTheta({0.9, 0.2, 0.5});
        c.SetIsMSPE(true);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        Locations<double> train(16, Dimension2D), test(2, Dimension2D);
        train.SetLocationX(*kX16, 16);
        train.SetLocationY(*kY16, 16);
        double tx[2] = {0.2, 0.3}, ty[2] = {0.2, 0.2};
        test.SetLocationX(*tx, 2);
        test.SetLocationY(*ty, 2);
        exageostat::api::ExaGeoStat<double>::ExaGeoStatPrediction(c, data, kZ16, &train, &test);
        auto zm = exageostat::results::Results::GetInstance()->GetPredictedMissedValues();
        REQUIRE(zm.size() == 2);
        REQUIRE(std::fabs(zm[0] - (-1.04437)) < 1e-5 && std::fabs(zm[1] - (-1.63133)) < 1e-5);
    });
    run("solver: ExaGeoStatMLETile at BOBYQA's first three points gives -24.026000 (TestExaGeoStatApi.cpp:93)", [] {
        Configurations c;
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(8);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto solver = exageostat::linearAlgebra::LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        using exageostat::kernels::Kernel;
        std::unique_ptr<Kernel<double>> k(
            exageostat::plugins::PluginRegistry<Kernel<double>>::Create("UnivariateMaternStationary", 1));
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        data->GetLocations()->SetLocationX(*kX16, 16);
        data->GetLocations()->SetLocationY(*kY16, 16);
        double z[16];
        std::memcpy(z, kZ16, sizeof(z));
        // lb, lb + (ub - lb)/4 e_1, lb + (ub - lb)/4 e_2 with lb = 0.1, ub = 5
        const double pts[3][3] = {{0.1, 0.1, 0.1}, {1.325, 0.1, 0.1}, {0.1, 1.325, 0.1}};
        double best = -INFINITY;
        for (auto &th : pts) best = std::max(best, solver->ExaGeoStatMLETile(data, c, th, z, *k));
        REQUIRE(std::fabs(best - (-24.026000)) < 1e-6);
        REQUIRE(data->GetMleIterations() == 3);
        // the observations were only read
        for (int i = 0; i < 16; ++i) REQUIRE(z[i] == kZ16[i]);
    });
    run("solver: GenerateObservationsVector reproduces z (TestExaGeoStatApi.cpp:52-54)", [] {
        Configurations c;
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(9);
        c.SetSeed(0);
        c.SetInitialTheta({1, 0.1, 0.5});
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto solver = exageostat::linearAlgebra::LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        using exageostat::kernels::Kernel;
        std::unique_ptr<Kernel<double>> k(
            exageostat::plugins::PluginRegistry<Kernel<double>>::Create("UnivariateMaternStationary", 1));
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        data->GetLocations()->SetLocationX(*kX16, 16);
        data->GetLocations()->SetLocationY(*kY16, 16);
        solver->GenerateObservationsVector(c, data, data->GetLocations(), data->GetLocations(), nullptr, 0, *k);
        for (int i = 0; i < 16; ++i) REQUIRE(std::fabs(data->GetMeasurements()[i] - kZ16[i]) < 1e-6);
    });
    run("solver: potrf, trsm and posv on the device descriptors", [] {
        Configurations c;
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(8);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto solver = exageostat::linearAlgebra::LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        void *seq = nullptr;
        solver->ExaGeoStatCreateSequence(&seq);
        REQUIRE(seq == ExaGeoStatHardware::GetContext());
        auto data = std::make_unique<ExaGeoStatData<double>>(16, Dimension2D);
        data->GetLocations()->SetLocationX(*kX16, 16);
        data->GetLocations()->SetLocationY(*kY16, 16);
        egs_problem_t p = data->GetProblem(ExaGeoStatHardware::GetContext(), EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY,
                                           0, 8);
        const double theta[3] = {1.0, 0.1, 0.5};
        // Sigma on the host through the plugin (nu = 1/2: exp(-d / range))
        std::vector<double> S(256);
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                const double dx = kX16[i] - kX16[j], dy = kY16[i] - kY16[j];
                S[i + 16 * j] = std::exp(-std::sqrt(dx * dx + dy * dy) / 0.1);
            }
        // potrf + trsm: ||L^-1 z||^2 equals z' Sigma^-1 z of an evaluation
        exageostat::CheckStatus(egs_covariance(p, theta), "egs_covariance");
        solver->ExaGeoStatPotrfTile(EXAGEOSTAT_LOWER, p, 0, nullptr, nullptr, 0, 0);
        std::vector<double> w(kZ16, kZ16 + 16);
        VectorDescriptor<double> dw{w.data(), 16};
        solver->ExaGeoStatTrsmTile(EXAGEOSTAT_LEFT, EXAGEOSTAT_LOWER, EXAGEOSTAT_NO_TRANS, EXAGEOSTAT_NON_UNIT, 1.0, p,
                                   nullptr, nullptr, &dw, 0);
        double q = 0.0;
        for (double v : w) q += v * v;
        // posv: Sigma x = z
        exageostat::CheckStatus(egs_covariance(p, theta), "egs_covariance");
        std::vector<double> x(kZ16, kZ16 + 16);
        VectorDescriptor<double> dx{x.data(), 16};
        solver->ExaGeoStatPosvTile(EXAGEOSTAT_LOWER, p, &dx);
        solver->ExaGeoStatSequenceWait(seq);
        double zx = 0.0;
        for (int i = 0; i < 16; ++i) {
            double r = 0.0;
            for (int j = 0; j < 16; ++j) r += S[i + 16 * j] * x[j];
            REQUIRE(std::fabs(r - kZ16[i]) < 1e-10);
            zx += kZ16[i] * x[i];
        }
        REQUIRE(std::fabs(q - zx) < 1e-10 * std::fabs(zx));
        // an upper factor is not what the tiles hold
        bool threw = false;
        try {
            solver->ExaGeoStatPotrfTile(EXAGEOSTAT_UPPER, p, 0, nullptr, nullptr, 0, 0);
        } catch (const std::domain_error &) {
            threw = true;
        }
        REQUIRE(threw);
    });
    run("solver: prediction returns {mspe, mspe1, mspe2} (LinearAlgebraMethods.cpp:596-600)", [] {
        Configurations c;
        c.SetKernelName("BivariateMaternParsimonious");
        c.SetDenseTileSize(8);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        auto solver = exageostat::linearAlgebra::LinearAlgebraFactory<double>::CreateLinearAlgebraSolver(EXACT_DENSE);
        using exageostat::kernels::Kernel;
        std::unique_ptr<Kernel<double>> k(
            exageostat::plugins::PluginRegistry<Kernel<double>>::Create("BivariateMaternParsimonious", 1));
        Locations<double> obs(12, Dimension2D), miss(4, Dimension2D);
        obs.SetLocationX(*kX16, 12);
        obs.SetLocationY(*kY16, 12);
        miss.SetLocationX(kX16[12], 4);
        miss.SetLocationY(kY16[12], 4);
        std::vector<double> zobs(24), zact(8), zmiss(8);
        for (int i = 0; i < 24; ++i) zobs[i] = kZ16[i % 16] * (i % 2 ? 0.5 : 1.0);
        for (int i = 0; i < 8; ++i) zact[i] = 0.1 * i - 0.3;
        double theta[6] = {1.0, 1.0, 0.1, 0.5, 1.0, 0.5};
        std::unique_ptr<ExaGeoStatData<double>> none;
        std::unique_ptr<double[]> e(solver->ExaGeoStatMLEPredictTile(none, theta, 4, 12, zobs.data(), zact.data(),
                                                                      zmiss.data(), c, miss, obs, *k));
        double s = 0.0, s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < 8; ++i) {
            const double d = (zact[i] - zmiss[i]) * (zact[i] - zmiss[i]);
            s += d;
            (i % 2 ? s2 : s1) += d;
        }
        REQUIRE(std::fabs(e[0] - s / 8) < 1e-14 && std::fabs(e[1] - s1 / 4) < 1e-14 && std::fabs(e[2] - s2 / 4) < 1e-14);
        REQUIRE(std::fabs(e[0] - 0.5 * (e[1] + e[2])) < 1e-14);
    });
    run("ExaGeoStatGenerateData is the synthetic branch of ExaGeoStatLoadData", [] {
        Configurations c;
        c.SetSeed(0);
        c.SetProblemSize(16);
        c.SetKernelName("UnivariateMaternStationary");
        c.SetDenseTileSize(9);
        c.SetInitialTheta({1, 0.1, 0.5});
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        std::unique_ptr<ExaGeoStatData<double>> a, b;
        exageostat::api::ExaGeoStat<double>::ExaGeoStatGenerateData(c, a);
        exageostat::api::ExaGeoStat<double>::ExaGeoStatLoadData(c, b);
        for (int i = 0; i < 16; ++i) {
            REQUIRE(a->GetMeasurements()[i] == b->GetMeasurements()[i]);
            REQUIRE(std::fabs(a->GetMeasurements()[i] - kZ16[i]) < 1e-6);
        }
    });
    run("kernel plugin evaluates a tile on the GPU", [] {
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 0);
        using exageostat::kernels::Kernel;
        std::unique_ptr<Kernel<double>> k(
            exageostat::plugins::PluginRegistry<Kernel<double>>::Create("UnivariateMaternStationary", 1));
        Locations<double> l(16, Dimension2D);
        l.SetLocationX(*kX16, 16);
        l.SetLocationY(*kY16, 16);
        double theta[3] = {1.0, 0.1, 0.5}, A[16 * 16];
        k->GenerateCovarianceMatrix(A, 16, 16, 0, 0, l, l, l, theta, 0);
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                const double dx = kX16[i] - kX16[j], dy = kY16[i] - kY16[j];
                const double want = std::exp(-std::sqrt(dx * dx + dy * dy) / 0.1);  // nu = 1/2 closed form
                REQUIRE(std::fabs(A[i + 16 * j] - want) < 1e-14);
            }
    });
}

static void mle_case(double t0, double t1, double t2, double llh_opt) {
    Configurations::SetVerbosity(QUIET_MODE);
    run("n=1600 fit from the lower bounds matches the oracle optimum", [&] {
        Configurations c;
        c.SetSeed(0);
        c.SetProblemSize(1600);
        c.SetDenseTileSize(320);
        c.SetInitialTheta({1, 0.1, 0.5});
        c.SetLowerBounds({0.01, 0.01, 0.01});
        c.SetUpperBounds({5, 5, 5});
        c.SetMaxMleIterations(1000);
        c.SetTolerance(9);
        auto hardware = ExaGeoStatHardware(EXACT_DENSE, 1, 1);
        std::unique_ptr<ExaGeoStatData<double>> data;
        exageostat::api::ExaGeoStat<double>::ExaGeoStatLoadData(c, data);
        const double llh = exageostat::api::ExaGeoStat<double>::ExaGeoStatDataModeling(c, data);
        auto &th = c.GetEstimatedTheta();
        std::printf("  theta-hat %.9f %.9f %.9f llh %.10f (%d evaluations)\n", th[0], th[1], th[2], llh,
                    data->GetMleIterations());
        REQUIRE(std::fabs(th[0] - t0) < 1e-4 && std::fabs(th[1] - t1) < 1e-4 && std::fabs(th[2] - t2) < 1e-4);
        REQUIRE(llh >= llh_opt - 1e-6);
    });
}

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") cpu_cases();
    else if (mode == "gpu") gpu_cases();
    else if (mode == "mle" && argc == 6) mle_case(std::atof(argv[2]), std::atof(argv[3]), std::atof(argv[4]), std::atof(argv[5]));
    else {
        std::fprintf(stderr, "usage: %s cpu | gpu | mle t0 t1 t2 llh\n", argv[0]);
        return 2;
    }
    std::printf("%s: %d failure(s)\n", mode.c_str(), g_failures);
    return g_failures == 0 ? 0 : 1;
}
