// Global configuration, same keys as the reference's conf:: namespace
// (src/utility/Configuration.hpp:14-119). fp_type is fixed to double: the
// plain path is FP64; the mixed-precision solve has its own entry points.
#pragma once

#include <cstddef>
#include <string>

namespace conf {
typedef double fp_type;

struct Configuration {
    int matrixBlockSize = 128;
    int workGroupSize = 128;
    int workGroupSizeVector = 128;
    int workGroupSizeFinalScalarProduct = 512;
    std::size_t iMax = 100000;
    double epsilon = 1.0e-6;
    int updateInterval = 10;
    double initialProportionGPU = 0.5;
    std::string outputPath = "./output";
    bool writeResult = false;
    bool writeMatrix = false;
    std::string mode = "static";
    std::size_t N = 0;
    std::size_t N_test = 0;
    double runtimeLBFactorCPU = 1.2;
    double energyLBFactorCPU = 0.8;
    std::size_t blockUpdateThreshold = 1;
    double idleWatt_CPU = 30.0;
    bool enableHWS = false;
    std::string algorithm = "cg";
    int workGroupSizeGEMM_xy = 16;
    int minBlockCountCholesky = 3;
    int gpuOptimizationLevel = 3;
    int cpuOptimizationLevel = 2;
    bool printVerbose = false;
    bool checkResult = false;
    bool trackCholeskySolveStep = true;
    // Keys of this build (no reference counterpart): CG, Cholesky and the
    // triangular solves in the reference's own floating-point order (its GPU
    // kernels' summation order, the work-group sizes above), so that results
    // compare with the reference's GPU build bit for bit; and whether that
    // order contracts multiply-adds.
    bool referenceOrder = false;
    bool referenceOrderFMA = true;
};

Configuration& get();

// The reference keeps matrixBlockSize as a by-value copy (Configuration.hpp:60);
// it is a plain variable here too.
inline unsigned long matrixBlockSize = 128;
inline int& workGroupSize = get().workGroupSize;
inline int& workGroupSizeVector = get().workGroupSizeVector;
inline int& workGroupSizeFinalScalarProduct = get().workGroupSizeFinalScalarProduct;
inline std::size_t& iMax = get().iMax;
inline double& epsilon = get().epsilon;
inline int& updateInterval = get().updateInterval;
inline double& initialProportionGPU = get().initialProportionGPU;
inline std::string& outputPath = get().outputPath;
inline bool& writeResult = get().writeResult;
inline bool& writeMatrix = get().writeMatrix;
inline std::string& mode = get().mode;
inline std::size_t& N = get().N;
inline std::size_t& N_test = get().N_test;
inline double& runtimeLBFactorCPU = get().runtimeLBFactorCPU;
inline double& energyLBFactorCPU = get().energyLBFactorCPU;
inline std::size_t& blockUpdateThreshold = get().blockUpdateThreshold;
inline double& idleWatt_CPU = get().idleWatt_CPU;
inline int& workGroupSizeGEMM_xy = get().workGroupSizeGEMM_xy;
inline int& minBlockCountCholesky = get().minBlockCountCholesky;
inline std::string& algorithm = get().algorithm;
inline bool& enableHWS = get().enableHWS;
inline int& gpuOptimizationLevel = get().gpuOptimizationLevel;
inline int& cpuOptimizationLevel = get().cpuOptimizationLevel;
inline bool& printVerbose = get().printVerbose;
inline bool& checkResult = get().checkResult;
inline bool& trackCholeskySolveStep = get().trackCholeskySolveStep;
inline bool& referenceOrder = get().referenceOrder;
inline bool& referenceOrderFMA = get().referenceOrderFMA;
}  // namespace conf
