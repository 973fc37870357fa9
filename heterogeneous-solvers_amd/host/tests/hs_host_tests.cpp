// Probe program for the C++ host API, driven by tests/test_host_api.py.
// Subcommands print what the pytest side compares with the reference's
// golden values (no GPU needed for the first six):
//   layout <matrix.txt> <bsz> <idx>...   values of A.matrixData at idx
//   shape <matrix.txt> <bsz>             blockSize blockCountXY size
//   rowparse <string>                    parsed values or invalid_argument
//   rhs <vector.txt> <bsz>               padded right-hand side
//   lb <share>                           StaticLoadBalancer validation
//   cgdims                               CG dimension checks
//   gpcpu                                matrixVectorGP on a CPU queue (rejected)
//   writefull <matrix.txt> <bsz> <out>   MatrixParser::writeFullMatrix of the parsed matrix
//   writeblocked <matrix.txt> <bsz> <out> MatrixParser::writeBlockedMatrix of it
//   metrics <algorithm> <dir> <count>    MetricsTracker::writeJSON with <count> entries per vector
//   solve <cg|cholesky> <share> <matrix.txt> <vector.txt> <bsz> <eps>
//                                        the solver classes at a GPU share (0: CPU only,
//                                        no device touched); prints iterations / info, then x
// and, on a GPU:
//   gemvgp <m> <n>                       matrixVectorGP on exact binary data
//   kstar <train> <test> <N> <M>         generateTestKernelMatrix, all entries
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "hs/hs.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::string cmd = argv[1];
    std::cout << std::setprecision(17);
    hs::Queue host(hs::cpu_selector_v);
    try {
        if (cmd == "layout" || cmd == "shape") {
            std::string path = argv[2];
            conf::matrixBlockSize = std::stoul(argv[3]);
            SymmetricMatrix A = MatrixParser::parseSymmetricMatrix(path, host);
            if (cmd == "shape") {
                std::cout << A.blockSize << " " << A.blockCountXY << " " << A.matrixData.size() << std::endl;
                return 0;
            }
            for (int i = 4; i < argc; ++i) std::cout << A.matrixData.at(std::stoul(argv[i])) << "\n";
            return 0;
        }
        if (cmd == "rowparse") {
            try {
                for (double v : MatrixParser::getRowValuesFromString(argv[2])) std::cout << v << "\n";
            } catch (const std::invalid_argument&) {
                std::cout << "invalid_argument" << std::endl;
            }
            return 0;
        }
        if (cmd == "rhs") {
            std::string path = argv[2];
            conf::matrixBlockSize = std::stoul(argv[3]);
            RightHandSide b = MatrixParser::parseRightHandSide(path, host);
            for (double v : b.rightHandSideData) std::cout << v << "\n";
            return 0;
        }
        if (cmd == "lb") {
            try {
                StaticLoadBalancer lb(10, std::stod(argv[2]), 4);
                MetricsTracker m;
                std::cout << "ok " << lb.getNewProportionGPU(m) << std::endl;
            } catch (const std::runtime_error& e) {
                std::cout << "runtime_error " << e.what() << std::endl;
            }
            return 0;
        }
        if (cmd == "cgdims") {
            conf::matrixBlockSize = 4;
            SymmetricMatrix A(20, 4, host);
            RightHandSide b(21, 4, host);
            RightHandSide c(20, 5, host);
            auto lb = std::make_shared<StaticLoadBalancer>(10, 1.0, A.blockCountXY);
            for (RightHandSide* r : {&b, &c}) {
                try {
                    CG cg(A, *r, host, host, lb);
                    std::cout << "no error" << std::endl;
                } catch (const std::invalid_argument& e) {
                    std::cout << "invalid_argument " << e.what() << std::endl;
                }
            }
            return 0;
        }
        if (cmd == "gpcpu") {
            double A[4] = {1, 2, 3, 4}, x[2] = {1, 1}, y[2];
            try {
                MatrixVectorOperations::matrixVectorGP(host, A, x, y, 2, 2);
                std::cout << "no error" << std::endl;
            } catch (const std::invalid_argument& e) {
                std::cout << "invalid_argument " << e.what() << std::endl;
            }
            return 0;
        }
        if (cmd == "writefull" || cmd == "writeblocked") {
            std::string path = argv[2];
            conf::matrixBlockSize = std::stoul(argv[3]);
            SymmetricMatrix A = MatrixParser::parseSymmetricMatrix(path, host);
            if (cmd == "writefull") MatrixParser::writeFullMatrix(argv[4], A);
            else MatrixParser::writeBlockedMatrix(argv[4], A);
            std::cout << "ok" << std::endl;
            return 0;
        }
        if (cmd == "metrics") {
            conf::algorithm = argv[2];
            std::string dir = argv[3];
            const int count = std::stoi(argv[4]);
            MetricsTracker m;
            for (int i = 0; i < count; ++i) {
                const double t = 0.5 + i;
                m.updateMetrics(i, 4, 0, t, conf::updateInterval);
                for (auto* v : {&m.times_q, &m.times_alpha, &m.times_x, &m.times_r, &m.times_delta, &m.times_d,
                                &m.memcopy_d, &m.matrixVectorTimes_GPU, &m.matrixVectorTimes_CPU, &m.shiftTimes,
                                &m.choleskyDiagonalBlockTimes, &m.copyTimes, &m.triangularSolveTimes_GPU,
                                &m.triangularSolveTimes_CPU, &m.triangularSolveTimes_total,
                                &m.matrixMatrixDiagonalTimes_GPU, &m.matrixMatrixDiagonalTimes_CPU,
                                &m.matrixMatrixDiagonalTimes_total, &m.matrixMatrixTimes_GPU,
                                &m.matrixMatrixTimes_CPU, &m.matrixMatrixTimes_total})
                    v->push_back(t / 4);
            }
            m.totalTime = 12.25;
            m.writeJSON(dir);
            std::cout << "ok" << std::endl;
            return 0;
        }
        if (cmd == "solve") {
            const std::string alg = argv[2];
            const double share = std::stod(argv[3]);
            std::string pathA = argv[4], pathB = argv[5];
            conf::matrixBlockSize = std::stoul(argv[6]);
            conf::workGroupSize = static_cast<int>(conf::matrixBlockSize);
            conf::epsilon = std::stod(argv[7]);
            conf::initialProportionGPU = share;
            conf::algorithm = alg;
            conf::outputPath = argc > 8 ? argv[8] : "./output";
            std::unique_ptr<hs::Queue> gpu;
            if (share > 0.0) gpu.reset(new hs::Queue(hs::gpu_selector_v));
            hs::Queue& gq = gpu ? *gpu : host;
            SymmetricMatrix A = MatrixParser::parseSymmetricMatrix(pathA, host);
            RightHandSide b = MatrixParser::parseRightHandSide(pathB, host);
            conf::N = A.N;
            auto lb = std::make_shared<StaticLoadBalancer>(conf::updateInterval, share, A.blockCountXY);
            std::vector<double> x;
            if (alg == "cg") {
                CG cg(A, b, host, gq, lb);
                cg.solveHeterogeneous();
                std::cout << "iterations " << cg.iterations << "\n";
                x.assign(cg.x.begin(), cg.x.end());
            } else {
                Cholesky chol(A, host, gq, lb);
                chol.solve_heterogeneous();
                TriangularSystemSolver tss(A, chol.A_gpu, b, host, gq, lb);
                tss.solve();
                std::cout << "info " << chol.info << "\n";
                x.assign(b.rightHandSideData.begin(), b.rightHandSideData.end());
            }
            for (double v : x) std::cout << v << "\n";
            return 0;
        }
        if (cmd == "gemvgp") {
            // entries are small multiples of 1/8 and 1/4: every product and
            // partial sum is exact, so any summation order gives the same bits
            const int m = std::stoi(argv[2]), n = std::stoi(argv[3]);
            std::vector<double> A((std::size_t)m * n), x(n), y(m);
            for (int r = 0; r < m; ++r)
                for (int k = 0; k < n; ++k) A[(std::size_t)r * n + k] = ((r * 7 + k * 3) % 11 - 5) * 0.125;
            for (int k = 0; k < n; ++k) x[k] = (k % 13 - 6) * 0.25;
            hs::Queue gpu(hs::gpu_selector_v);
            MatrixVectorOperations::matrixVectorGP(gpu, A.data(), x.data(), y.data(), n, m);
            for (double v : y) std::cout << v << "\n";
            return 0;
        }
        if (cmd == "kstar") {
            std::string train = argv[2], test = argv[3];
            conf::N = std::stoul(argv[4]);
            conf::N_test = std::stoul(argv[5]);
            std::vector<double> K(conf::N * conf::N_test);
            hs::Queue gpu(hs::gpu_selector_v);
            MatrixGenerator::generateTestKernelMatrix(train, test, host, gpu, K.data());
            for (double v : K) std::cout << v << "\n";
            return 0;
        }
    } catch (const std::exception& e) {
        std::cout << "error " << e.what() << std::endl;
        return 1;
    }
    return 2;
}
