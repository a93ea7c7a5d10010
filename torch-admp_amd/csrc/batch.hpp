// Batched PME engine: many independent periodic systems of different sizes,
// boxes and meshes evaluated by one launch sequence (see batch.hip).
#pragma once

#include <vector>

#include "admp_common.hpp"
#include "nblist_kernels.hpp"
#include "pme_kernels.hpp"
#include "tiles.hpp"

namespace admp {

constexpr int kGatherAtomsPerBlock = kBlock / 8;
// gather partial values per block: the dE/dA terms of the reciprocal sum and
// of sum r (x) dE/dr over the real-space gradient rows (in one 3 x 3 form)
constexpr int kGatherNV = 9;
constexpr int kSysValues = kRealNV + 10 + kGatherNV;  // reduced values per system

// Per-system description, built on the host and read by every batched kernel.
struct SysDesc {
  Geom g;
  CellGrid cg;
  TileGeom tg;
  Radix rad[3];
  double kappa, eps;
  int conv, ortho;
  int fchunk[3];  // lines per block: z passes (row pairs), y passes, x pass
  int plane;      // the y and x passes run as one LDS-resident plane per kz
                  // (the half spectrum is then stored kz-major)
  int brick;      // spread by mesh bricks (owner computes, no atomics): 1 bricks of
                  // kBrickZ points along z, kBrickWaves deeper bricks (dense meshes)
  int nbr[3];     // bricks per mesh axis
  int nsplit;     // real-space candidate splits (work-groups) per home cell
  int kzp;        // row pitch of the half spectrum (line layout: K2/2 + 1 rounded up to
                  // whole 128-byte lines; plane layout: K2/2 + 1)
  int64_t atom0, n;
  int64_t cell0, ncell;
  int64_t mesh0, spec0;
};

// Segment offsets (n_sys + 1 each) used to map a global index to its system.
// kSegFftX numbers the x-pass partial rows of every system (Kzh planes or the
// line work-groups); kSegPlane and kSegColsX number the work-groups of the
// plane pass and of the x line pass only, so that neither kernel is launched
// over the other's systems.
enum SegKind {
  kSegAtom = 0, kSegCell, kSegMesh, kSegFftZ, kSegFftY, kSegFftX, kSegGather, kSegBrick,
  kSegPlane, kSegColsX, kSegBrickZ, kSegReal, kSegCount
};

struct BatchPlan {
  int nsys;
  int precision;
  std::vector<SysDesc> sys;
  std::vector<int64_t> off[kSegCount];
  int64_t ntot, ctot, mtot, stot;
  size_t lds_fft[3];   // dynamic LDS of the z, y and x passes
  size_t lds_tile;
  size_t lds_plane;    // dynamic LDS of the plane pass
  int any_plane, any_cols;  // some system takes the plane pass / the line passes
  int any_untiled;     // some system spreads with global atomics
  int any_brick;       // some system spreads by mesh bricks
  int any_nonbrick;    // some system spreads into a zeroed mesh (tiles or atomics)
  int nsplit;          // fewest real-space candidate splits per cell (each system
                       // takes at least this many, SysDesc::nsplit)
  NbLayout nb;         // binning workspace over all cells
  CellGrid cg_all;     // only ncell is used by the shared binning helpers
  // workspace byte offsets
  size_t o_tab, o_off, o_nb, o_sys_of, o_ps, o_mref, o_stm, o_stw, o_grad, o_mesh, o_spec, o_rpart,
      o_cpart, o_gpart, o_red, o_tickets, bytes;
  int status;          // ADMP_OK or an error code from planning
};

// Host planning of a batch; status != ADMP_OK when a system is unsupported.
BatchPlan plan_batch(int nsys, const int64_t* n_atoms, const double* boxes, const double* cutoffs,
                     const double* kappas, const int32_t* kmesh, int precision, int convention,
                     double eps);

// Runs the whole launch sequence.  Outputs: energy (nsys x 5: real,
// reciprocal, self, non-neutral, total), forces (ntot x 3), dEdbox (nsys x 9, may be null).
// `upload` = false assumes upload_table already put this plan's table in ws
// (single systems never need it: the first kernel writes their table).
// `fork` = false keeps every kernel on `s` (no side stream is created or used).
int run_batch(const BatchPlan& plan, const double* pos, const double* q, double* energy,
              double* forces, double* dEdbox, char* ws, hipStream_t s, hipEvent_t* marks,
              bool upload = true, bool fork = true);

// The pieces of a batch that a charge solve repeats.  batch_geometry: table,
// binning and cell order of the positions (q: any per-atom values, packed
// beside them).  batch_potential: out[i] += the reciprocal-space potential
// dE_recip/dq_i of the charges v (spread, FFT and influence function, gather),
// on the geometry of the last batch_geometry.
int batch_geometry(const BatchPlan& plan, const double* pos, const double* q, char* ws,
                   hipStream_t s);
int batch_potential(const BatchPlan& plan, const double* pos, const double* v, double* out,
                    char* ws, hipStream_t s);

// Copy the system table and segment offsets into the workspace; `sync`
// waits for the copy.
int upload_table(const BatchPlan& plan, char* ws, hipStream_t s, bool sync);

}  // namespace admp
