// tuning.hpp -- the autotuned GEMM kernel table.
//
// Each entry names, for one precision mode and one GEMM shape class
// (log2 M, log2 N, log2 K, log2 batch, rounded), the kernel family and tile
// variant that ran it fastest when jet_hip_autotune_gemm timed every valid
// candidate (jet_amd/autotune.py sweeps the shape grid on the device and
// writes jet_amd/data/gemm_tuning_gfx950.json, which the build compiles
// into the library).  A plan takes the entry of its own shape class, else
// the nearest entry of its mode (L1 distance in log2 coordinates, at most
// kTuneMaxDistance) whose choice is valid for the shape, else the built-in
// rule (choose_gemm_kernel).
#pragma once

#include "gemm.hpp"

#include <cstdint>
#include <vector>

namespace jet_hip {

struct TuneEntry {
    int32_t mode, lm, ln, lk, lb; // precision mode, log2 M / N / K / batch
    int32_t kernel, variant;      // the fastest choice
};

constexpr int kTuneEntryInts = 7;
constexpr double kTuneMaxDistance = 4.0;

// Replaces the process-wide table (thread-safe).
void tuning_set(const std::vector<TuneEntry> &entries);
// Restores the table compiled into the library.
void tuning_reset();
std::vector<TuneEntry> tuning_get();

// source (may be null): 0 built-in rule, 1 the shape's own entry, 2 the
// nearest entry.
GemmChoice tuned_gemm_choice(int mode, int64_t M, int64_t N, int64_t K, int64_t batch,
                             int *source);

} // namespace jet_hip
