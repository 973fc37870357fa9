// Command-line front end with the reference's flags (src/main.cpp:37-187) and
// the same flow (main.cpp:189-252): parse or generate A and b, build the load
// balancer, run CG or Cholesky + triangular solves (or, with --gpr, the
// Gaussian-process prediction after the solve), write metrics.json.
// Flags take the form --name=value or --name value.
#include <cmath>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <string>

#include "hs/hs.hpp"

static bool toBool(const std::string& v) { return v == "1" || v == "true" || v == "True" || v.empty(); }

int main(int argc, char** argv) {
    std::map<std::string, std::string> a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        if (s.rfind("--", 0) != 0) {
            std::cerr << "unexpected argument " << s << std::endl;
            return 2;
        }
        s = s.substr(2);
        const auto eq = s.find('=');
        if (eq != std::string::npos) a[s.substr(0, eq)] = s.substr(eq + 1);
        else if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) a[s] = argv[++i];
        else a[s] = "";
    }
    try {
        std::string pathA = a.count("path_A") ? a["path_A"] : "";
        std::string pathB = a.count("path_b") ? a["path_b"] : "";
        std::string gpIn = a.count("gp_input") ? a["gp_input"] : "";
        std::string gpOut = a.count("gp_output") ? a["gp_output"] : "";
        if (a.count("output")) conf::outputPath = a["output"];
        if (a.count("mode")) conf::mode = a["mode"];
        if (a.count("matrix_bsz")) {
            conf::matrixBlockSize = std::stoul(a["matrix_bsz"]);
            conf::workGroupSize = static_cast<int>(conf::matrixBlockSize);
        }
        if (a.count("i_max")) conf::iMax = std::stoul(a["i_max"]);
        if (a.count("eps")) conf::epsilon = std::stod(a["eps"]);
        if (a.count("update_int")) conf::updateInterval = std::stoi(a["update_int"]);
        if (a.count("init_gpu_perc")) conf::initialProportionGPU = std::stod(a["init_gpu_perc"]);
        if (a.count("write_result")) conf::writeResult = toBool(a["write_result"]);
        if (a.count("write_matrix")) conf::writeMatrix = toBool(a["write_matrix"]);
        if (a.count("size")) conf::N = std::stoul(a["size"]);
        if (a.count("algorithm")) conf::algorithm = a["algorithm"];
        if (a.count("print_verbose")) conf::printVerbose = toBool(a["print_verbose"]);
        if (a.count("check_result")) conf::checkResult = toBool(a["check_result"]);
        if (a.count("test_size")) conf::N_test = std::stoul(a["test_size"]);
        // this build's flags: CG in the reference's floating-point order
        if (a.count("reference_order")) conf::referenceOrder = toBool(a["reference_order"]);
        if (a.count("fma")) conf::referenceOrderFMA = toBool(a["fma"]);
        std::string gpTest = a.count("gp_test") ? a["gp_test"] : "";
        const bool performGPR = a.count("gpr") && toBool(a["gpr"]);
        const bool generate = !gpIn.empty();

        hs::Queue gpuQueue(hs::gpu_selector_v);
        hs::Queue cpuQueue(hs::cpu_selector_v);
        std::cout << "GPU: " << gpuQueue.name() << std::endl;

        RightHandSide b = generate ? MatrixGenerator::parseRHS_GP(gpOut, gpuQueue)
                                   : MatrixParser::parseRightHandSide(pathB, gpuQueue);
        SymmetricMatrix A = generate ? MatrixGenerator::generateSPDMatrix(gpIn, cpuQueue, gpuQueue)
                                     : MatrixParser::parseSymmetricMatrix(pathA, gpuQueue);
        if (conf::mode != "static") {
            throw std::runtime_error("Invalid mode selected: '" + conf::mode +
                                     "' --> this build has the 'static' load balancer only");
        }
        auto lb = std::make_shared<StaticLoadBalancer>(conf::updateInterval, conf::initialProportionGPU,
                                                       A.blockCountXY);
        std::vector<double> b0(b.rightHandSideData.begin(), b.rightHandSideData.end());
        std::vector<double> xres;
        if (performGPR) {
            // main.cpp:222-224: solve, then predict at the test inputs
            if (!generate) throw std::invalid_argument("--gpr needs --gp_input and --gp_output");
            GaussianProcess GP(A, b, gpIn, gpTest, cpuQueue, gpuQueue, lb);
            GP.start();
            double s = 0.0;
            for (double v : GP.result) s += v;
            std::cout << "RESULT algorithm=gpr test_size=" << GP.result.size() << " mean_sum=" << std::setprecision(17)
                      << s << std::endl;
            return 0;
        }
        if (conf::algorithm == "cg") {
            CG cg(A, b, cpuQueue, gpuQueue, lb);
            cg.solveHeterogeneous();
            xres.assign(cg.x.begin(), cg.x.end());
            std::cout << "RESULT algorithm=cg iterations=" << cg.iterations << " delta_ratio="
                      << (cg.deltaZero > 0 ? cg.deltaFinal / cg.deltaZero : 0.0) << std::endl;
        } else if (conf::algorithm == "cholesky") {
            SymmetricMatrix Acopy = A;  // keep A for the residual check
            Cholesky chol(A, cpuQueue, gpuQueue, lb);
            chol.solve_heterogeneous();
            TriangularSystemSolver solver(A, chol.A_gpu, b, cpuQueue, gpuQueue, lb);
            const double ms = solver.solve();
            chol.metricsTracker.solveTime = ms;
            chol.writeMetricsToFile();
            xres.assign(b.rightHandSideData.begin(), b.rightHandSideData.end());
            std::cout << "RESULT algorithm=cholesky info=" << chol.info << std::endl;
            A.matrixData = Acopy.matrixData;
        } else {
            throw std::runtime_error("Invalid algorithm: " + conf::algorithm);
        }
        if (conf::checkResult) {
            // mean |A x - b| as UtilityFunctions::checkResult (UtilityFunctions.cpp:70-91), mat-vec on the GPU
            hs_handle_t h = gpuQueue.handle();
            gpuQueue.reserve(A.N, A.blockSize);
            hs::DeviceArray<double> dA(A.matrixData.size()), dx(xres.size()), dy(xres.size());
            hs::check(hs_memcpy_htod(h, dA.data(), A.matrixData.data(), A.matrixData.size() * 8), "htod");
            hs::check(hs_memcpy_htod(h, dx.data(), xres.data(), xres.size() * 8), "htod");
            hs::check(hs_symv_packed_f64(h, dA.data(), dx.data(), dy.data(), static_cast<int64_t>(A.N), A.blockSize),
                      "hs_symv_packed_f64");
            std::vector<double> y(xres.size());
            hs::check(hs_memcpy_dtoh(h, y.data(), dy.data(), y.size() * 8), "dtoh");
            double err = 0.0, rn = 0.0, bn = 0.0;
            for (std::size_t i = 0; i < A.N; ++i) {
                err += std::fabs(y[i] - b0[i]);
                rn += (y[i] - b0[i]) * (y[i] - b0[i]);
                bn += b0[i] * b0[i];
            }
            std::cout << "Average error of Ax - b: " << err / static_cast<double>(A.N) << std::endl;
            std::cout << "RESULT relative_residual=" << std::sqrt(rn / bn) << std::endl;
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
