// Device view of a walker ensemble: model tables + per-walker state, all in
// HBM.  Passed by value to every kernel.
#pragma once
#include <stdint.h>

namespace mvmc_hip {

constexpr int NFLAGS = 8;  // ints of an ensemble's walker counters (EnsView::nflag)

struct EnsView {
  // dimensions
  int W;          // walkers
  int n, ne, ns, n2;  // Nsize, Ne, Nsite, 2*Nsite
  int nqp, nleg, nmp;
  int nproj, ngutz, nslater, size;  // size = 1 + NPara
  int nsample, warmup, interval, ex_path;
  int ntrans, nintra, ninter, nhund, nexch;
  int orb_injective;  // each orbital-index row is injective
  int exp_cr;         // 1: correctly rounded exp/log (always on the device)
  // model tables
  const double* slater;   // [nqp][n2][n2]
  const double* qpw;      // [nqp]
  const double* proj;     // [nproj]
  const int* gutz_idx;    // [ns]
  const int* jast_idx;    // [ns*ns] symmetric, -1 on the diagonal
  const int* loc_spn;     // [ns]
  const int* trans;       // [ntrans][4] site1 spin1 site2 spin2
  const double* trans_v;  // [ntrans]
  const int* intra_site;  const double* intra_v;
  const int* inter_site;  const double* inter_v;  // [ninter][2]
  const int* hund_site;   const double* hund_v;
  const int* exch_site;   const double* exch_v;
  const double* dproj_g;  // [ns]
  const double* dproj_j;  // [ns*ns]
  const int* tsite;       // [nmp][ns]
  const int* tsgn;        // [nmp][ns]
  const int* orb_idx;     // [ns*ns]
  const int* orb_sgn;     // [ns*ns]
  const double* spgl_cs; const double* spgl_cc; const double* spgl_ss;  // [nleg]
  // walker state
  uint32_t* rng;      // [W][625]: 624 words + cursor
  int* ele_idx;       // [W][n]
  int* ele_cfg;       // [W][n2]
  int* ele_num;       // [W][n2]
  int* proj_cnt;      // [W][nproj]
  double* inv;        // [W][nqp][n*n]
  float* inv32;       // [W][nqp][n*n] the sampler's inverse in the mixed-precision form, else null
  double* pf;         // [W][nqp]
  double* log_ip;     // [W]
  int* wstate;        // [W][8]: 0 burn flag, 1 n_accept window, 2 n_out, 3 status,
                      //   4 next proposal of the outer step, 5 rebuild pending, 6 outer step done,
                      //   7 left by k_begin_rr to k_begin
  int* nflag;         // [NFLAGS]: 0 walkers handed to the rebuild kernel by the last k_block launch,
                      //   1 next walker taken by a k_block slot, 2 next hand-over entry taken by k_rebuild,
                      //   3 passes of the last k_block launch that wrote a walker's planes back,
                      //   4 those of them that accepted no move (pending updates flushed),
                      //   5 walkers k_begin_rr left to k_begin
  int* rlist;         // [W] walkers handed to the rebuild kernel, in hand-over order
  long long* counter; // [W][4]
  // saved samples
  int* s_idx;  int* s_cfg;  int* s_num;  int* s_proj;  // [W][nsample][...]
  // measurement outputs
  double* e_loc;      // [W][nsample]
  int* s_valid;       // [W][nsample]
  double* sz;         // [W][nsample]
  double* o_store;    // [W][nsample][size]
  // per-slot scratch (persistent grids: slot = blockIdx.x)
  double* work;       // [slots][nqp][n*n]
  double* qorb;       // [slots][nqp][nslater]
  double* pend;       // [slots][nqp][n][PEND_MAX][2]: vec1 / vec2 of the hop updates not yet written back
  int* status;        // [1] sticky
  long long* clk;     // [slots][32] phase clocks (MVMC_PHASE_CLOCKS builds), else null
  double* gscr;       // [slots][gstride] per-term Green ratios of the measurement
  int gstride;
};

// One launch over several ensembles (runs of different models): run r owns
// workgroups [slot_start[r], slot_start[r+1]) of the grid, which serve its
// walkers exactly as a launch of its own would (slot = blockIdx.x -
// slot_start[r], its own scratch and walker counters).
struct GroupArgs {
  const EnsView* views;   // [nruns] device copies of the runs' views
  const int* slot_start;  // [nruns + 1]
  int nruns;              // 0: single-ensemble launch (the view argument)
  long long* slot_time;   // [grid] busy clock ticks per workgroup (accumulated), or null
  // Shared walker queue of the sweep and rebuild launches: queue entry g lies
  // in [qstart[k], qstart[k+1]) for run qrun[k], walker g - qstart[k]; any
  // workgroup takes any run's walker, with scratch slot blockIdx.x of the
  // group's own buffers (strides for the largest run).
  int* gq;                // [4]: 0 entries handed to the rebuild, 1 next entry, 2 next hand-over taken
  int* glist;             // [total] queue entries handed to the rebuild
  const int* qstart;      // [nruns + 1]
  const int* qrun;        // [nruns]
  double* work;           // [grid][work_stride]
  double* pend;           // [grid][pend_stride]
  long long work_stride, pend_stride;
};

}  // namespace mvmc_hip
