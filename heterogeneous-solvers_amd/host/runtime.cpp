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

template <typename T>
static std::string toJSON(const std::vector<T>& v) {
    std::ostringstream s;
    s << std::setprecision(17) << "[";
    for (std::size_t i = 0; i < v.size(); ++i) s << (i ? "," : "") << v[i];
    s << "]";
    return s.str();
}

void MetricsTracker::writeJSON(std::string& path) {
    std::ofstream f(path + "/metrics.json");
    f << "{\n\"configuration\": {\n";
    f << "\t \"algorithm\": \"" << conf::algorithm << "\",\n";
    f << "\t \"N\": " << conf::N << ",\n";
    f << "\t \"FP_type\": \"FP_64\",\n";
    f << "\t \"matrixBlockSize\": " << conf::matrixBlockSize << ",\n";
    f << "\t \"workGroupSize\": " << conf::workGroupSize << ",\n";
    f << "\t \"workGroupSizeVector\": " << conf::workGroupSizeVector << ",\n";
    f << "\t \"workGroupSizeFinalScalarProduct\": " << conf::workGroupSizeFinalScalarProduct << ",\n";
    f << "\t \"iMax\": " << conf::iMax << ",\n";
    f << "\t \"epsilon\": " << conf::epsilon << ",\n";
    f << "\t \"updateInterval\": " << conf::updateInterval << ",\n";
    f << "\t \"initialProportionGPU\": " << conf::initialProportionGPU << ",\n";
    f << "\t \"runtimeLBFactorCPU\": " << conf::runtimeLBFactorCPU << ",\n";
    f << "\t \"blockUpdateThreshold\": " << conf::blockUpdateThreshold << ",\n";
    f << "\t \"gpuOptimizationLevel\": " << conf::gpuOptimizationLevel << ",\n";
    f << "\t \"cpuOptimizationLevel\": " << conf::cpuOptimizationLevel << ",\n";
    f << "\t \"enableHWS\": " << conf::enableHWS << ",\n";
    const char* bind = std::getenv("OMP_PROC_BIND");
    const char* cores = std::getenv("OMP_NUM_THREADS");
    f << "\t \"cpuBind\": \"" << (bind ? bind : "") << "\",\n";
    f << "\t \"ompCores\": " << (cores ? cores : "\"\"") << ",\n";
    f << "\t \"mode\": \"" << conf::mode << "\"\n";
    f << "},\n\"runtime-metrics\": {\n";
    f << "\t \"iterationTimes\": " << toJSON(iterationTimes) << ",\n";
    f << "\t \"memoryInitTime\": " << memoryInitTime << ",\n";
    f << "\t \"resultCopyTime\": " << resultCopyTime << ",\n";
    f << "\t \"totalTime\": " << totalTime << ",\n";
    f << "\t \"choleskySolveStepTime\": " << solveTime << ",\n";
    f << "\t \"averageUtilization_GPU\": " << toJSON(averageUtilization_GPU) << ",\n";
    f << "\t \"averageUtilization_CPU\": " << toJSON(averageUtilization_CPU) << ",\n";
    f << "\t \"powerDraw_GPU\": " << toJSON(powerDraw_GPU) << ",\n";
    f << "\t \"powerDraw_CPU\": " << toJSON(powerDraw_CPU) << ",\n";
    f << "\t \"blockCounts_GPU\": " << toJSON(blockCounts_GPU) << ",\n";
    f << "\t \"blockCounts_CPU\": " << toJSON(blockCounts_CPU) << ",\n";
    f << "\t \"matrixVectorTimes_GPU\": " << toJSON(matrixVectorTimes_GPU) << ",\n";
    f << "\t \"matrixVectorTimes_CPU\": " << toJSON(matrixVectorTimes_CPU) << ",\n";
    f << "\t \"times_q\": " << toJSON(times_q) << ",\n";
    f << "\t \"times_alpha\": " << toJSON(times_alpha) << ",\n";
    f << "\t \"times_x\": " << toJSON(times_x) << ",\n";
    f << "\t \"times_r\": " << toJSON(times_r) << ",\n";
    f << "\t \"times_delta\": " << toJSON(times_delta) << ",\n";
    f << "\t \"times_d\": " << toJSON(times_d) << ",\n";
    f << "\t \"memcopy_d\": " << toJSON(memcopy_d) << ",\n";
    f << "\t \"shiftTimes\": " << toJSON(shiftTimes) << ",\n";
    f << "\t \"choleskyDiagonalBlockTimes\": " << toJSON(choleskyDiagonalBlockTimes) << ",\n";
    f << "\t \"copyTimes\": " << toJSON(copyTimes) << ",\n";
    f << "\t \"triangularSolveTimes_GPU\": " << toJSON(triangularSolveTimes_GPU) << ",\n";
    f << "\t \"triangularSolveTimes_CPU\": " << toJSON(triangularSolveTimes_CPU) << ",\n";
    f << "\t \"triangularSolveTimes_total\": " << toJSON(triangularSolveTimes_total) << ",\n";
    f << "\t \"matrixMatrixDiagonalTimes_GPU\": " << toJSON(matrixMatrixDiagonalTimes_GPU) << ",\n";
    f << "\t \"matrixMatrixDiagonalTimes_CPU\": " << toJSON(matrixMatrixDiagonalTimes_CPU) << ",\n";
    f << "\t \"matrixMatrixDiagonalTimes_total\": " << toJSON(matrixMatrixDiagonalTimes_total) << ",\n";
    f << "\t \"matrixMatrixTimes_GPU\": " << toJSON(matrixMatrixTimes_GPU) << ",\n";
    f << "\t \"matrixMatrixTimes_CPU\": " << toJSON(matrixMatrixTimes_CPU) << ",\n";
    f << "\t \"matrixMatrixTimes_total\": " << toJSON(matrixMatrixTimes_total) << "\n";
    f << "}\n}\n";
}
