// Bin-sort geometry and entry points (k_bin.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace migrid {

// Bins of H pixels over q = floor(coordinate) + S/2, q in [0, N+S-2].
// A pixel x receives contributions from visibilities with q in [x, x+S-1].
struct BinGeom {
    int N = 0, S = 0, H = 0;
    int nb = 0;         // bins per axis
    int64_t nkeys = 0;  // 64 * nb * nb
};

BinGeom bin_geom(int64_t N, int64_t S, int H);

// One job of a multi-job sort (device-resident table).  Its elements are
// [e_base, e_base + npts) of the concatenated arrays; its keys are
// [k_base, k_base + nkeys], the last one the sentinel of visibilities that
// touch no pixel.
struct BinJob {
    const void* uv;
    int64_t npts;
    int64_t e_base;
    int64_t k_base;
    int N, S, H, nb;
};
// Writes n host job descriptions to device memory (kernel launches; graph-safe).
int put_binjobs(BinJob* dst, const BinJob* src, int n, hipStream_t s);

// Scratch bytes for E elements over K keys.
size_t bin_scratch_bytes(int64_t E, int64_t K);

// Stable sort of the E elements of njobs jobs by key (K keys in all, K <=
// 2^31): perm[0..E) the element indices in key order (each job's binned
// visibilities, then its sentinels, job after job); offsets[0..K) the CSR
// starts, offsets[k] = number of elements with key < k -- for job j,
// offsets[k_base + nkeys] ends its binned visibilities.  Asynchronous.
int bin_sort(const BinJob* jobs, int njobs, int64_t max_npts, int64_t E, int64_t K, int in_type,
             void* scratch, int32_t* perm, int32_t* offsets, hipStream_t s);

// The same sort for keys the caller has written (E keys, each < K) into
// sort_key_buffer(scratch); scratch as bin_scratch_bytes(E, K).
uint32_t* sort_key_buffer(void* scratch);
int sort_keys(int64_t E, int64_t K, void* scratch, int32_t* perm, int32_t* offsets, hipStream_t s);

// Groups npts visibilities by a GCF index in [0, ngcf) (k_inputs.hip):
// uv_out/vals_out receive them in group order, stable within a group, and
// offsets[0..ngcf] the group starts (offsets[ngcf] starts the indexes out of
// range).  scratch: bin_scratch_bytes(npts, ngcf + 1); perm: npts ints.
int group_by_index(const int32_t* idx, int64_t npts, int ngcf, int in_type, const void* uv,
                   const void* vals, void* uv_out, void* vals_out, void* scratch, int32_t* perm,
                   int32_t* offsets, hipStream_t s);

int64_t scan_scratch_elems(int64_t n);
int exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* partial, bool write_total,
                   hipStream_t s);

}  // namespace migrid
