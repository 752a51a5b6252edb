// batched.hpp -- device records and host driver of the batched solver.
#pragma once

#include <vector>

#include "internal.hpp"
#include "solver.hpp"

namespace nosh {

// Metadata of one padded slot for the FP32 row product, 12 bytes: internal
// column, alpha and a~ signed by direction (alpha 0 for padding).
struct SlotMeta {
  int32_t col;
  float alpha, at;
};
static_assert(sizeof(SlotMeta) == 12, "slot metadata is 12 packed bytes");

// The FP32 off-diagonal value of one padded slot at a given mu, 12 bytes:
// internal column and K = vr + i vi = -alpha e^{-i mu a~}, i.e.
// vr = -alpha cos(mu a~), vi = alpha sin(mu a~) (the same FP32 operations as
// the row product's own rebuild, so either path gives the same bits).
struct SlotVal {
  int32_t col;
  float vr, vi;
};
static_assert(sizeof(SlotVal) == 12, "slot values are 12 packed bytes");

// One mesh and the instances solved on it (device-side record).
struct BGroup {
  int64_t n;        // rows (vertices)
  int bp;           // lanes per row: instances padded to a power of two
  int B;            // instances
  int inst0;        // first global instance index
  int64_t block0;   // first block of this group in a batched launch
  int64_t nblocks;  // blocks of this group
  const int32_t* slice_ptr;
  const int32_t* scol;
  const SlotMeta* s32;  // per padded slot, FP32 path
  const SlotVal* kval;  // per padded slot at the mu every active instance of the
                        // group shares, or null (each lane rebuilds K at its mu)
  const double2* kval64;  // per padded slot, FP64 value -alpha e^{-i mu a~} at the mu
                          // every instance of the group shares, or null (F then
                          // rebuilds it per lane)
  const double* a64;
  const double* t64;
  const double* diag;    // per row: sum alpha
  const double* rowabs;  // per row: sum |alpha|
  const double* ct;      // per row: control volume * thickness
  const double* V;       // per row: scalar potential
  double2* psi;          // [n][bp], FP64 state
  double2* f;            // [n][bp], F(psi)
  double2* x;            // [n][bp], MINRES solution accumulator
  float* jblk;           // 3 [n][bp]
  float* pblk;           // 3 [n][bp]
  float2 *Y, *Z2, *v, *u, *R0, *R1, *W0, *W1, *W2;  // [n][bp] FP32 Krylov vectors
  float2* v1;  // second Lanczos buffer: v of iteration k lives in (k odd ? v1 : v)
};

struct BatchArgs {
  const BGroup* groups;
  int ngroups;
  int ninst;
  int64_t total_blocks;  // virtual blocks (block_threads / bp rows each)
  int64_t maxblocks;
  int fold;              // virtual blocks per launched block
  MinresScalars* S;     // per instance
  const double* mu;     // per instance
  const double* g;      // per instance
  double* bound;        // per instance: max ratio for the Chebyshev interval
  const double* rtol;   // per instance MINRES tolerance
  const int* active;    // per instance: 1 while Newton runs
  const double* shift;  // per instance: pseudo-transient shift sigma (J + sigma c t I)
  double upper, ratio;
  int iters;            // Chebyshev iterations
  int z0;               // parity of the first Chebyshev output buffer
  double* partials;     // [instance][maxblocks]
  unsigned* bar;        // persistent mode: grid barrier {arrivals, generation}
  int* bar_err;         // persistent mode: set when a barrier wait times out
  long long bar_ticks;  // persistent mode: barrier wait limit in wall-clock ticks
};

struct GroupHost {
  MeshCache* mesh = nullptr;
  int B = 0;
  DBuf<SlotMeta> s32;
  DBuf<SlotVal> kval;     // allocated on first use
  double kval_mu = 0.0;   // the mu kval holds (valid when kval_set)
  bool kval_set = false;
  DBuf<double2> kval64;   // FP64 slot values for F, allocated on first use
  double kval64_mu = 0.0;
  bool kval64_set = false;
  DBuf<double> a64, t64, diag, rowabs, ct, V;
  DBuf<double2> psi, f, x;
  DBuf<float> jblk, pblk;
  DBuf<float2> Y, Z2, v, v1, u, R0, R1, W0, W1, W2;
  BGroup dev{};
};

// Workgroups per CU of the persistent (cooperative) kernel's grid: at most
// two, and only as many as its registers allow. The kernel strides over all row
// tiles, so any grid size is correct as long as every workgroup is resident
// at once; a grid sized to the occupancy limit (occupancy x CUs) leaves no
// margin, since the dispatcher hands workgroups to the XCDs round-robin and
// every CU then has to hold exactly its share.
constexpr int kPersistentPerCu = 2;
// Wait limit of one grid barrier of the persistent kernel (seconds). A wait
// that exceeds it (workgroups not co-resident) makes every workgroup leave
// the kernel and the solve report an error instead of hanging.
constexpr double kBarrierTimeoutS = 2.0;

// Lanes per thread of the gathering kernels in tuning mode 2.
constexpr int kVecLanes = 2;

struct Batch {
  std::vector<GroupHost> ghost;
  DBuf<BGroup> groups;
  BatchArgs args{};
  DBuf<MinresScalars> S;
  std::vector<MinresScalars> Sh;
  DBuf<double> mu, g, bound, rtol, norms, partials, shift;
  DBuf<int> active;
  DBuf<unsigned> bar;  // persistent mode: grid barrier state
  DBuf<int> bar_err;   // persistent mode: barrier time-out flag
  std::vector<int> active_h;
  int block_threads = 256;
  int fold = 1;         // virtual blocks per launched block (fewer reduction partials)
  int mode = 0;         // 0: hipGraph-captured iterations, 1: persistent cooperative kernel
  int lanes = 1;        // lanes per thread of the gathering kernels in hipGraph mode
                        // (1 or kVecLanes)
  int graph_iters = 4;  // MINRES iterations per captured graph
  int poll = 16;        // iterations between host checks of the done flags
  hipGraphExec_t graph_exec = nullptr;
  std::vector<double> mu_host;      // per instance, the values uploaded to mu
  std::vector<BGroup> records_host;  // staging of the device group records

  void relayout(int threads, int fold);
  // Uploads per-instance mu (and keeps the host copy).
  void set_mu(const double* m, hipStream_t s);
  // For every group whose active instances all share one mu (and whose rows
  // take slot metadata by lane shuffles, or with lanes > 1), fills kval at
  // that mu (once per value) and points the group record at it; other groups
  // rebuild K per lane.
  void prepare_kval(hipStream_t s);
  // For every group whose instances all share one mu, fills kval64 at that
  // mu (once per value) for the FP64 residual; other groups rebuild K per
  // lane there.
  void prepare_kval64(hipStream_t s, bool* all = nullptr);
  void upload_records(hipStream_t s);
  void launch_phases_iteration(hipStream_t s);
  void launch_cheb_step(hipStream_t s, int mode, int j);
  void launch_minres_a(hipStream_t s);
  void build_graph(hipStream_t s);
  void drop_graph();
  int persistent_grid();
  // Rebuilds blocks at psi, solves J dx = -F for the active instances and
  // applies psi += dx; returns the largest MINRES iteration count.
  int minres(hipStream_t s, int maxit);
  void compute_norms(hipStream_t s, std::vector<double>& nrm);
  // f := scale * dF/dmu at psi and each instance's mu, for every instance
  // (then minres() gives psi + scale * dpsi/dmu: the tangent predictor).
  void dfdmu_into_f(hipStream_t s, double scale);
  // Average duration (us) of one Chebyshev step launch (kb_cheb_step,
  // j = min(2, iters - 1)) with every instance active, timed with events on s over reps launches
  // after rebuilding at the current psi. Leaves the MINRES state reset.
  double time_cheb_step(hipStream_t s, int reps);
  ~Batch();
};

// Per-slot (alpha, signed a~) in FP32 and FP64 plus row sums, from the edge
// weights of the mesh (operators.hip).
void launch_slot_fields(const MeshCache& m, const double* alpha, const double* atilde,
                        SlotMeta* s32, double* a64, double* t64, double* diag, double* rowabs,
                        hipStream_t s);
// Reference-order instance vector -> lane b of an interleaved [n][bp] array, and back.
void launch_scatter_lane(int64_t n, int bp, int b, const int32_t* perm, const double* src,
                         double* dst, hipStream_t s);
void launch_gather_lane(int64_t n, int bp, int b, const int32_t* perm, const double* src,
                        double* dst, hipStream_t s);

}  // namespace nosh
