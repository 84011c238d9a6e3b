// Pseudo-ECG on the hex block (ecg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cg_internal.hpp"

namespace cg {

constexpr int ECG_BLOCKS = 1024;       // most partial sums per electrode
constexpr int ECG_MAX_ELECTRODES = 256;

// Allocation-free pseudo-ECG (cg_tissue_pseudo_ecg_noalloc): at most this many electrodes,
// one kernel pass.  Its scratch is a file-scope device array of ecg.hip, one per GPU:
// [3 * ECG_NOALLOC_ELECTRODES] positions, [ECG_NOALLOC_ELECTRODES] results, then
// [ECG_NOALLOC_ELECTRODES][ECG_BLOCKS] per-block partial sums.
constexpr int ECG_NOALLOC_ELECTRODES = 16;
constexpr int ECG_SCRATCH_DOUBLES = ECG_NOALLOC_ELECTRODES * (3 + 1 + ECG_BLOCKS);
// device address of the current GPU's scratch (looked up once per device with
// hipGetSymbolAddress); 0 or a hipError_t
hipError_t ecg_scratch(double** out);

// blocks of the first kernel for n cells (<= ECG_BLOCKS): partial holds [E][that]
int ecg_partial_blocks(int64_t n);

// out_dev[e] = pseudo-ECG at electrode e from V (device, blockMesh order) and the
// conductivity: sig per cell (device; [6][n] xx xy xz yy yz zz when sym, else [9][n]) or,
// when null, the uniform s9 (host [9]).  pos_dev [E][3] device; partial [E][ECG_BLOCKS]
// device scratch.  Asynchronous on s.
hipError_t launch_pseudo_ecg(const HexGeom& g, const double* V, const double* sig, int sym,
                             const double* s9, int n_el, const double* pos_dev,
                             double* partial, double* out_dev, hipStream_t s);

} // namespace cg
