// SCC launch choices of single-point batches by basis size on MI355X, written
// by qcxms_amd/autotune.py: {max_nao, MFMA eigen refinement, LDS matrices}
// per bin, and {eig, lds} for batches spanning several bins (eig -1: one
// launch per run of systems of one bin).
#pragma once

namespace qx {

constexpr int kTuneTable[][3] = {
    {40, 1, 1},
    {56, 0, 1},
    {256, 1, 0},
};
constexpr int kTuneMixed[2] = {1, 1};

}  // namespace qx
