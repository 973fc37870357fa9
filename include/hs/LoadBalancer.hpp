// GPU-share policy interface (src/loadBalancer/LoadBalancer.hpp:10-28) and the
// static policy (StaticLoadBalancer.cpp:6-14). The solvers run the whole
// solve on the GPU at share 1.0 and split rows between CPU and GPU below it.
#pragma once

#include <stdexcept>

#include "hs/MetricsTracker.hpp"

class LoadBalancer {
public:
    LoadBalancer(int updateInterval, double initialProportionGPU, int blockCountXY)
        : updateInterval(updateInterval), currentProportionGPU(initialProportionGPU), blockCountXY(blockCountXY) {}
    virtual double getNewProportionGPU(MetricsTracker& metricsTracker) = 0;
    int updateInterval;
    double currentProportionGPU;
    int blockCountXY;
    virtual ~LoadBalancer() {}
};

class StaticLoadBalancer : public LoadBalancer {
public:
    StaticLoadBalancer(int updateInterval, double gpuProportion, int blockCountXY)
        : LoadBalancer(updateInterval, gpuProportion, blockCountXY), gpuProportion(gpuProportion) {
        if (gpuProportion > 1.0 || gpuProportion < 0.0) {
            throw std::runtime_error("Invalid GPU proportion. Must be between 0 and 1");
        }
    }
    double getNewProportionGPU(MetricsTracker&) override { return gpuProportion; }
    double gpuProportion;
};
