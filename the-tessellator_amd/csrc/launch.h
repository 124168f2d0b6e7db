// launch.h -- host-side launchers of the device passes (one per kernel file).
#pragma once
#include "tess_internal.h"

namespace tess {

// bin.hip
hipError_t launch_soa(const double* xyz, int n, double* x, double* y, double* z, hipStream_t s);
hipError_t launch_bounds(const double* x, const double* y, const double* z, int n, int cpd,
                         double* part, int nparts, GridDev* g, hipStream_t s);
// counting sort: bins/cnt/delim/cursor/sorted/scratch sized by the caller
struct SortBufs {
  uint32_t* bins;     // [n]
  uint32_t* cnt;      // [nbins]
  uint32_t* delim;    // [nbins + 1]
  uint32_t* cursor;   // [nbins]
  uint32_t* sorted;   // [n]
  uint32_t* scratch;  // [n]
  uint32_t* sums;     // [ceil(nbins / 1024) + 1]
  uint32_t* big;      // [nbins]
  uint32_t* misc;     // [4]: total, nbig
};
hipError_t launch_counting_sort(const double* x, const double* y, const double* z, int n,
                                const GridDev* g, int nbins, const SortBufs& b, hipStream_t s);
hipError_t launch_gather(const uint32_t* sorted, int n, const double* x, const double* y,
                         const double* z, double* sx, double* sy, double* sz, hipStream_t s);
hipError_t launch_face_offsets(const uint32_t* fcount, int n, uint64_t* foff, uint64_t* sums,
                               hipStream_t s);
hipError_t launch_compact(const uint32_t* fcount, const uint64_t* foff, int n, const int32_t* snbr,
                          const double* sarea, int32_t* nbr, double* area, hipStream_t s);
hipError_t launch_invert(const uint32_t* sorted, int n, uint32_t* inv, hipStream_t s);

// cell.hip: vpl = vertex slots per lane (1, 2 or 4)
hipError_t launch_cells(int vpl, const CellArgs& a, hipStream_t s);
hipError_t launch_diag_arith(const double* a, const double* b, int n, double* sq, double* rc,
                             double* dv, hipStream_t s);

}  // namespace tess
