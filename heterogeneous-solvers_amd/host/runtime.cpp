// Queue, data structures and metrics of the C++ host API.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "hs/hs.hpp"

namespace conf {
Configuration& get() {
    static Configuration c;
    return c;
}
}  // namespace conf

namespace hs {

void check(int status, const char* what) {
    if (status < 0) {
        const char* msg = hs_last_error_string();
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(status) + "): " + (msg ? msg : ""));
    }
}

bool pinned_host_available() {
    static const bool available = [] {
        int count = 0;
        return hs_device_count(&count) == HS_OK && count > 0;
    }();
    return available;
}

Queue::Queue(gpu_selector_t, int device) : device_(device) {
    check(hs_handle_create(&handle_, device), "hs_handle_create");
}

Queue::Queue(cpu_selector_t) {}

Queue::~Queue() {
    if (handle_) hs_handle_destroy(handle_);
}

hs_handle_t Queue::handle() const {
    if (!handle_) throw std::runtime_error("this queue is a host (CPU) queue: it has no device");
    return handle_;
}

void Queue::wait() {
    if (handle_) check(hs_sync(handle_), "hs_sync");
}

void Queue::reserve(std::size_t N, int b, std::size_t M) {
    check(hs_workspace_reserve(handle(), static_cast<int64_t>(N), b, static_cast<int64_t>(M)),
          "hs_workspace_reserve");
}

std::size_t Queue::global_mem_size() const {
    if (!handle_) return 0;
    size_t fr = 0, tot = 0;
    check(hs_device_memory(handle_, &fr, &tot), "hs_device_memory");
    return tot;
}

std::string Queue::name() const { return handle_ ? "MI355X (gfx950) device " + std::to_string(device_) : "host"; }

}  // namespace hs

SymmetricMatrix::SymmetricMatrix(std::size_t N, int blockSize, hs::Queue&)
    : N(N),
      blockSize(blockSize),
      blockCountXY(static_cast<int>(std::ceil(static_cast<double>(N) / static_cast<double>(blockSize)))) {
    const std::size_t blockCount = static_cast<std::size_t>(blockCountXY) * (blockCountXY + 1) / 2;
    matrixData.assign(blockCount * static_cast<std::size_t>(blockSize) * blockSize, 0.0);
}

RightHandSide::RightHandSide(std::size_t N, int blockSize, hs::Queue&)
    : N(N),
      blockSize(blockSize),
      blockCountX(static_cast<int>(std::ceil(static_cast<double>(N) / static_cast<double>(blockSize)))) {
    rightHandSideData.assign(static_cast<std::size_t>(blockCountX) * blockSize, 0.0);
}

void MetricsTracker::startTracking() {}
void MetricsTracker::endTracking() {}

void MetricsTracker::updateMetrics(std::size_t, std::size_t blockCount_GPU, std::size_t blockCount_CPU,
                                   double iterationTime, int) {
    blockCounts_GPU.push_back(blockCount_GPU);
    blockCounts_CPU.push_back(blockCount_CPU);
    iterationTimes.push_back(iterationTime);
}

// vectorToJSONString (MetricsTracker.cpp:309-325): "[a, b, c]" with each
// value through std::to_string (fixed, six decimals)
template <typename T>
static std::string toJSON(const std::vector<T>& v) {
    std::string s = "[";
    for (std::size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i]);
    return s + "]";
}

// MetricsTracker::writeJSON (MetricsTracker.cpp:152-307) as built without the
// hws sampler: the configuration block, the common runtime metrics, then the
// CG or the Cholesky vectors; scalars through std::to_string.
void MetricsTracker::writeJSON(std::string& path) {
    std::ofstream f(path + "/metrics.json");
    auto kv = [&](const char* key, const std::string& value, bool last = false) {
        f << "\t \"" << key << "\": " << value << (last ? "\n" : ",\n");
    };
    auto str = [](const std::string& v) { return "\"" + v + "\""; };
    f << "{\n\"configuration\": {\n";
    kv("algorithm", str(conf::algorithm));
    kv("N", std::to_string(conf::N));
    kv("FP_type", str("FP_64"));
    kv("matrixBlockSize", std::to_string(conf::matrixBlockSize));
    kv("workGroupSize", std::to_string(conf::workGroupSize));
    kv("workGroupSizeVector", std::to_string(conf::workGroupSizeVector));
    kv("workGroupSizeFinalScalarProduct", std::to_string(conf::workGroupSizeFinalScalarProduct));
    kv("iMax", std::to_string(conf::iMax));
    kv("epsilon", std::to_string(conf::epsilon));
    kv("updateInterval", std::to_string(conf::updateInterval));
    kv("initialProportionGPU", std::to_string(conf::initialProportionGPU));
    kv("runtimeLBFactorCPU", std::to_string(conf::runtimeLBFactorCPU));
    kv("blockUpdateThreshold", std::to_string(conf::blockUpdateThreshold));
    kv("gpuOptimizationLevel", std::to_string(conf::gpuOptimizationLevel));
    kv("cpuOptimizationLevel", std::to_string(conf::cpuOptimizationLevel));
    kv("enableHWS", std::to_string(conf::enableHWS));
    const char* bind = std::getenv("OMP_PROC_BIND");
    const char* cores = std::getenv("OMP_NUM_THREADS");
    kv("cpuBind", str(bind ? bind : ""));
    kv("ompCores", cores ? std::string(cores) : std::string("\"\""));
    kv("mode", str(conf::mode), true);
    f << "},\n\"runtime-metrics\": {\n";
    kv("iterationTimes", toJSON(iterationTimes));
    kv("memoryInitTime", std::to_string(memoryInitTime));
    kv("resultCopyTime", std::to_string(resultCopyTime));
    kv("totalTime", std::to_string(totalTime));
    kv("choleskySolveStepTime", std::to_string(solveTime));
    kv("averageUtilization_GPU", toJSON(averageUtilization_GPU));
    kv("averageUtilization_CPU", toJSON(averageUtilization_CPU));
    kv("powerDraw_GPU", toJSON(powerDraw_GPU));
    kv("powerDraw_CPU", toJSON(powerDraw_CPU));
    kv("blockCounts_GPU", toJSON(blockCounts_GPU));
    const bool cg = conf::algorithm == "cg", chol = conf::algorithm == "cholesky";
    kv("blockCounts_CPU", toJSON(blockCounts_CPU), !cg && !chol);
    if (cg) {
        kv("matrixVectorTimes_GPU", toJSON(matrixVectorTimes_GPU));
        kv("matrixVectorTimes_CPU", toJSON(matrixVectorTimes_CPU));
        kv("times_q", toJSON(times_q));
        kv("times_alpha", toJSON(times_alpha));
        kv("times_x", toJSON(times_x));
        kv("times_r", toJSON(times_r));
        kv("times_delta", toJSON(times_delta));
        kv("times_d", toJSON(times_d));
        kv("memcopy_d", toJSON(memcopy_d), true);
    } else if (chol) {
        kv("shiftTimes", toJSON(shiftTimes));
        kv("choleskyDiagonalBlockTimes", toJSON(choleskyDiagonalBlockTimes));
        kv("copyTimes", toJSON(copyTimes));
        kv("triangularSolveTimes_GPU", toJSON(triangularSolveTimes_GPU));
        kv("triangularSolveTimes_CPU", toJSON(triangularSolveTimes_CPU));
        kv("triangularSolveTimes_total", toJSON(triangularSolveTimes_total));
        kv("matrixMatrixDiagonalTimes_GPU", toJSON(matrixMatrixDiagonalTimes_GPU));
        kv("matrixMatrixDiagonalTimes_CPU", toJSON(matrixMatrixDiagonalTimes_CPU));
        kv("matrixMatrixDiagonalTimes_total", toJSON(matrixMatrixDiagonalTimes_total));
        kv("matrixMatrixTimes_GPU", toJSON(matrixMatrixTimes_GPU));
        kv("matrixMatrixTimes_CPU", toJSON(matrixMatrixTimes_CPU));
        kv("matrixMatrixTimes_total", toJSON(matrixMatrixTimes_total), true);
    }
    f << "}\n}\n";
}
