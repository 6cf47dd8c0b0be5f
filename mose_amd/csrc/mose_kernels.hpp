// mose_kernels.hpp — device data structures and kernels of the explicit step (gfx950).
//
// Memory layout in HBM
//   * Cell pool: the padded cells (gc = 2 ghost layers) of every block concatenated,
//     i fastest.  Conserved state is SoA over the pool: U[v * npool + cell], v = rho, rho u,
//     rho v, rho w, rho E, FP64.  Three state buffers: U^n, and two stage buffers.
//   * Geometry of rectilinear blocks: per block and direction a line of `LineEnt`
//     (cell length dl as the reference computes it, node spacing, reconstruction spacings
//     of the face on the cell's low side).  Normals are the axis unit vectors.
//   * Geometry of curvilinear blocks (any block not rectilinear switches the context to it):
//     per pool cell, SoA over the pool, NGEO FP64 arrays: the unit normal and area of the
//     cell's low face in each direction, the cell lengths dl, the volume, the unit rows of
//     the inverse Jacobian (CFL time step) and the reciprocal distance to the low neighbour's
//     centre (reconstruction), all from the reference's metric definitions.
//   * Work: a list of tiles (block, i0, j0, k-range).  Each workgroup of TX x TY threads
//     owns a TX x TY column patch and marches along k; the current k-plane's primitive
//     state with a 2-cell i/j halo is staged in LDS, the k-direction stencil of each column
//     lives in registers.  Every face flux is evaluated once.
#pragma once
#include "mose_device.hpp"

namespace mose {


// Ghost-op kinds.
enum { G_COPY = 0, G_ZG = 1, G_EXTRAP = 2, G_SYM = 3 };
// Face BC kinds used inside the stage kernel (bctype of a block face).
enum { F_CONN = 0, F_NONE = 1, F_EXTRAP = 2, F_SYM = 3, F_INFLOW = 4, F_INLET = 5, F_OUTFLOW = 6 };
// Scalars kept on the device.
enum { S_DT = 0, S_TIME = 1, S_NORM = 2, S_COUNT = 8 };
// Curvilinear geometry arrays (index g: array g * npool + cell).
//   GN + 4d + {0,1,2}: unit normal of the low face in direction d, GN + 4d + 3: its area;
//   GDL + d: dl; GVOL: volume; GVS + 3d + {0,1,2}: M(d,:)/|M(d,:)|; GRLO + d: 2/(dl(c-e_d)+dl(c)).
enum { GN = 0, GDL = 12, GVOL = 15, GVS = 16, GRLO = 25, NGEO = 28 };

struct LineEnt {
  double dl;      // cell length along the line (Compute_Metric_Tensor dl)
  double delta;   // node spacing x_i - x_{i-1} (face areas / volume)
  double rl0, rl1, rl2, lm, lp;   // spacing of the face between cells i-1 and i
  double rdl;     // 1 / dl (the diagonal metric of the viscous terms)
};

// Read-only data through the constant address space: loads from a wave-uniform address
// become scalar loads (s_load, scalar data cache) instead of vector loads into VGPRs.  Only
// for arrays no kernel writes while the loading kernel runs.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

struct BlockDev {
  int64_t off;       // pool index of padded cell (-1,-1,-1)
  int64_t sz;        // plane stride (ni+4)(nj+4)
  int ni, nj, nk, sy;
  int line[3];       // LineEnt index of padded position 0 (cell -1) per direction
  int bct[6];        // F_* per face
  int bcd[6];        // index of the 8-double data record (inflow), or -1
  int bcs[6];        // first slot of the face in the boundary-flux buffer, or -1
  int bvs[6];        // first slot of a connection face in the viscous-flux buffer, or -1
  int visc, pad_;    // laminar viscous terms on
  double vis[5];     // mu_ref, T_ref, S (Sutherland), cp / Pr, vnn
};

// One boundary face cell of a block connection when viscous terms are on: the stencil of
// BC_Connection_Visc (boundary cell, its ghost, the tangential neighbours of both sides).
struct ViscFace {
  int64_t loc, gh;     // pool index of the boundary cell and of its first ghost cell
  int64_t tn[4];       // boundary cell's neighbours: 1st tangential -, +, 2nd tangential -, +
  int64_t dn[4];       // the donor cell's neighbours in the same order (Pg(:,3:6))
  int b, d, slot, modfm2;
  double area;
  double ml[3], mg[3]; // diagonal metric (1/dl) of the boundary and ghost cells
};

// One boundary face cell of a non-connection face (flux computed by k_bcface).
struct BcFace {
  int64_t cb, cn;    // pool index of the boundary cell and of its interior neighbour
  int b, f;          // block (0-based), face (0..5)
  int slot, pad;     // output slot in the boundary-flux buffer
  double bdl, idl;   // half lengths of the two cells along the face normal
  double n[3];       // unit normal of the boundary face (increasing-index orientation)
  double area;       // face area
  double t[3];       // unit face diagonal (t_Vec of Setup_Inflow_Geometry, subsonic inlets)
};

struct Tile {
  int b, i0, j0, k0, k1, pad0, pad1, pad2;
};

// Largest patch (TX x TY threads) of any tile shape.  Every kernel launched with one thread
// per patch column (the stage kernels and k_dt, k_sd, k_prim_resid, k_irs_jacobi,
// k_irs_update) declares a launch bound of at least this, and the host refuses a shape above
// it (a launch with more threads than a kernel's bound fails).
constexpr int kMaxTileThreads = 512;

struct GhostOp {
  int64_t b;         // boundary cell (pool index)
  int64_t s;         // inward pool stride
  int64_t src1, src2;
  int type, pad;
  double n[3];
  // Archive slots of the op's second and third reads (copy: src1 / src2; extrapolation:
  // b + s / b + 2s), or -1: a read of a ghost cell that a later record of the same fill
  // writes takes that cell's value from the previous fill, as the reference's serial fill
  // does (blocks one or two cells thick).
  int a1, a2;
};

struct StageArgs {
  const double* __restrict__ Uin;
  const double* __restrict__ Un;
  double* __restrict__ Uout;
  double* __restrict__ Rout;
  const Tile* __restrict__ tiles;
  const BlockDev* __restrict__ blocks;
  const LineEnt* __restrict__ lines;
  const double* __restrict__ facebc;
  const double* __restrict__ geo;      // curvilinear geometry arrays (NGEO x npool) or null
  const double* __restrict__ scal;
  const double* dtloc;          // local time stepping: dt of each cell (same array as dtnext)
  double* __restrict__ partials;
  double* __restrict__ bcbuf;
  double* __restrict__ vbuf;    // viscous flux of connection faces (per unit area), viscous path
  const double* __restrict__ beta;   // shock-detector weight per pool cell (HLLE++ / HLLC+ / MUSCL-SD)
  int* err;
  unsigned long long* dtmin;   // last stage: CFL time step of the next step (min as bits)
  double* dtnext;              // last stage, local time stepping: dtlocal of the next step
  const int* iter;
  int* tq;                     // tile queues of the persistent stage launch (kTileQueues claim
                               // counters and a finished-workgroup counter), zero between launches
  int64_t npool;
  int ntiles;
  int nq;                      // queues in use: kTileQueues with the XCD-aware tile order, else 1
  int local_dt;
  int ndir, rampa;
  int irs_scale;               // MODE 2: write R* = -R dt / vol (residual smoothing) instead of R
  double rkc, gam, R, cfl;
};

// Persistent stage launch: the grid is the number of workgroups resident on the device and each
// workgroup claims tiles from a queue until every queue is empty.  Queue x (x < nq) holds the tile
// positions x, x + nq, x + 2 nq, ...: with the XCD-aware order those are the tiles of one
// contiguous range of blocks, and workgroup w (dispatched to XCD w mod 8) starts on queue w mod nq,
// so neighbouring tiles share an L2.  A workgroup whose queue is empty takes tiles from the next
// queues.  Every claim is one atomicAdd, so each tile is taken exactly once; the last workgroup to
// finish sets the counters back to zero for the next launch.
constexpr int kTileQueues = 8;

__device__ __forceinline__ int claim_tile(int* tq, int nq, int ntiles) {
  const int x0 = (int)(blockIdx.x % (unsigned)nq);
  for (int s = 0; s < nq; ++s) {
    const int x = x0 + s < nq ? x0 + s : x0 + s - nq;
    const int len = (ntiles - x + nq - 1) / nq;
    if (__hip_atomic_load(tq + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
    const int q = atomicAdd(tq + x, 1);
    if (q < len) return x + q * nq;
  }
  return -1;
}

// End of a workgroup's persistent loop (thread 0): the last workgroup resets the queues.
__device__ __forceinline__ void finish_tiles(int* tq) {
  if (atomicAdd(tq + kTileQueues, 1) == (int)gridDim.x - 1) {
    for (int x = 0; x <= kTileQueues; ++x) atomicExch(tq + x, 0);
  }
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads() also orders global memory,
// which makes every wave wait for all of its outstanding global loads before the barrier;
// kernels whose threads exchange data only through LDS use this one, so that loads issued
// ahead of time (the next plane's cells) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// First error of a launch wins (code, block, i, j, k); later ones are dropped.
__device__ __forceinline__ void set_error(int* err, int code, int b, int i, int j, int k) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = b; err[2] = i; err[3] = j; err[4] = k;
  }
}

}  // namespace mose
