// step.h -- launchers of the per-timestep kernels (step.hip).
#pragma once

#include "internal.h"

namespace uvlm {

void launch_geometry(hipStream_t s, const double* nodes, int num, int R, int C,
                     double* cp, double* nrm, double* area);
void launch_matrix(hipStream_t s, const RingRec* recs, const double* cp,
                   const double* nrm, int N, const CutOff& cut, double* A);
// Same matrix in the rsqrt form of the FP64-chain and F32C modes.
void launch_matrix_fast(hipStream_t s, const RingRec* recs, const double* cp,
                        const double* nrm, int N, const CutOff& cut, double* A);
// Legacy ring order (A[L*N + K] = ring_L.BiotSavartLaw(cp_K, 1) . n_K).
void launch_ring_matrix(hipStream_t s, const RingRec* recs, const double* cp,
                        const double* nrm, int N, const CutOff& cut, double* A);
void launch_rhs(hipStream_t s, const double* uw, const double* vm,
                const double* nrm, int N, const double* ff, double* rhs);
void launch_shed(hipStream_t s, const double* wing, const double* gamma, int num,
                 int R, int C, double* wake, int cap, int row, double* wgam,
                 int capg, int grow);
void launch_gather_rows(hipStream_t s, const double* rb, int cap, int row0,
                        int num, int rows, int C1, double* out);
void launch_stage_pos(hipStream_t s, const double* wake, const double* k,
                      double dt, double div, size_t total, double* out);
void launch_update(hipStream_t s, int scheme, const double* pos,
                   const double* k1, const double* k2, const double* k3,
                   const double* k4, const double* vprev, int prev_rows, int ign,
                   double dt, int num, int rows, int C1, double* rb, int cap,
                   int row0);
void launch_edge_midpoints(hipStream_t s, const double* nodes, int num, int R,
                           int C, double* mid);
void launch_loads(hipStream_t s, const double* nodes, const double* gamma,
                  const double* gamma_prev, const double* nrm,
                  const double* area, const double* uflow, const double* vm,
                  int num, int R, int C, double dt, double Q,
                  const double* ref_area, double* out);

}  // namespace uvlm
