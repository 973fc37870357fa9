// Timing record of a solve, same fields and metrics.json keys as the
// reference (src/metricsTracker/MetricsTracker.hpp:21-111,
// MetricsTracker.cpp:152-307). The hws power/utilisation sampling is not
// part of this build: those vectors stay empty.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

class MetricsTracker {
public:
    std::vector<double> averageUtilization_GPU;
    std::vector<double> averageUtilization_CPU;
    std::vector<double> powerDraw_GPU;
    std::vector<double> powerDraw_CPU;
    std::vector<std::size_t> blockCounts_GPU;
    std::vector<std::size_t> blockCounts_CPU;
    std::vector<double> matrixVectorTimes_GPU;
    std::vector<double> matrixVectorTimes_CPU;
    std::vector<double> times_q;
    std::vector<double> times_alpha;
    std::vector<double> times_x;
    std::vector<double> times_r;
    std::vector<double> times_delta;
    std::vector<double> times_d;
    std::vector<double> memcopy_d;
    std::vector<double> shiftTimes;
    std::vector<double> choleskyDiagonalBlockTimes;
    std::vector<double> copyTimes;
    std::vector<double> triangularSolveTimes_GPU;
    std::vector<double> triangularSolveTimes_CPU;
    std::vector<double> triangularSolveTimes_total;
    std::vector<double> matrixMatrixDiagonalTimes_GPU;
    std::vector<double> matrixMatrixDiagonalTimes_CPU;
    std::vector<double> matrixMatrixDiagonalTimes_total;
    std::vector<double> matrixMatrixTimes_GPU;
    std::vector<double> matrixMatrixTimes_CPU;
    std::vector<double> matrixMatrixTimes_total;
    std::vector<double> iterationTimes;

    double memoryInitTime = 0.0;
    double resultCopyTime = 0.0;
    double totalTime = 0.0;
    double solveTime = 0.0;

    void startTracking();
    void endTracking();
    void updateMetrics(std::size_t iteration, std::size_t blockCount_GPU, std::size_t blockCount_CPU,
                       double iterationTime, int updateInterval);
    void writeJSON(std::string& path);
};
