// Abstract force (simulation/redmax/Force/Force.h:9-58): the interface Robot's force
// loops call (computeForce / computeForceWithDerivative accumulate into the global
// maximal-coordinate fm, Km, Dm and reduced-coordinate fr, Kr, Dr) and the per-force
// timing report.
#pragma once
#include <cstdio>
#include <iostream>
#include <map>
#include <string>
#include "Common.h"
#include "Utils.h"

namespace redmax {

class Simulation;

class Force {
public:
    Simulation* _sim;

    // Accumulated durations per named phase, in clock() ticks (microseconds),
    // printed as one table row per phase in milliseconds.
    class Timer {
        std::map<std::string, long long> totals_;

    public:
        void add(std::string name, long long duration) { totals_[name] += duration; }
        void print_time_report() {
            for (const auto& kv : totals_) {
                char row[96];
                std::snprintf(row, sizeof row, "|%-20s|%7lld(ms) |", kv.first.c_str(), kv.second / 1000);
                std::cerr << row << std::endl;
            }
        }
        void clear() { totals_.clear(); }
        long long get(const std::string& name) const {
            const auto it = totals_.find(name);
            return it == totals_.end() ? 0 : it->second;
        }
    } _time;

    explicit Force(Simulation* sim) { _sim = sim; }
    virtual ~Force() = default;

    virtual void init() {}
    virtual void computeForce(VectorX& fm, VectorX& fr, bool verbose = false) = 0;
    virtual void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr, MatrixX& Dr,
                                            bool verbose = false) = 0;
    virtual void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr,
                                            MatrixX& Dr, MatrixX& dfm_dp, MatrixX& dfr_dp, bool verbose = false) {}

    virtual void test_derivatives_runtime() {}
    virtual void test_design_derivatives_runtime() {}

    void reset_time_report() { _time.clear(); }
    void print_time_report() { _time.print_time_report(); }
};

}  // namespace redmax
