// colphys_handle.h -- the column-physics handle (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "colphys_kernels.h"

namespace colphys {
struct SurrState;
struct MomState;
// momentum vertical viscosity (momentum.hip)
int mom_solve(rdb_colphys* h, double dt, const rdb_colphys_momentum_config* c);
void mom_release(rdb_colphys* h);
int mom_face_set(rdb_colphys* h, bool x_face);
int mom_face_get(rdb_colphys* h, bool x_face);
int mom_field_spec(rdb_colphys* h, const char* name, int* e1, int* e2, int* e3, bool* settable);
int mom_set_field(rdb_colphys* h, const char* name, const double* data, size_t n);
int mom_get_field(rdb_colphys* h, const char* name, double* out, size_t n);
bool handle_valid(rdb_colphys* h);
// shared by the C ABI translation units
int launch_step(rdb_colphys* h, double dt, int p0, int p1);
int refresh_static(rdb_colphys* h, double dt);
// n surrogate steps, captured in a hipGraph and replayed (surrogate.hip)
int surr_run(rdb_colphys* h, double dt, int n);
// a launch argument of the captured surrogate step changed (the state view)
void surr_invalidate(rdb_colphys* h);
// the handle's step event pair (created on first use)
int step_events(rdb_colphys* h, hipEvent_t* e0, hipEvent_t* e1);
// the last surrogate step's fallback count and forward time, read from the
// device (synchronises the stream when a step is pending)
int surr_settle(rdb_colphys* h);
void surr_free(SurrState* s);
int surr_collect(rdb_colphys* h, double dt, int n_steps, int stride, long long* n_out);
int surr_train(rdb_colphys* h, const rdb_colphys_mlp_config* cfg, float* train_loss, float* val_loss, int n_hist,
               int* n_done);
int surr_save(rdb_colphys* h, const char* path);
int surr_load(rdb_colphys* h, const char* path);
int surr_set_guard(rdb_colphys* h, int mode);
bool surr_ready(rdb_colphys* h);
}  // namespace colphys

int rdb_colphys_push_error(int status, const char* msg);

struct rdb_colphys {
  rdb_colphys_config cfg;
  colphys::Params pr;
  int nx, ny, nz, n_passive;
  int n_all, n_wet, P;
  std::vector<int> col_host;
  std::vector<unsigned char> dry_host;  // per full-grid column
  int* d_col = nullptr;
  colphys::StateView s{};
  double* d_stage = nullptr;
  size_t stage_elems = 0;
  double *d_txc = nullptr, *d_tyc = nullptr, *d_smag = nullptr, *d_shelf = nullptr;
  unsigned int* d_guard = nullptr;
  unsigned long long* d_acc = nullptr;
  hipStream_t stream = nullptr;
  bool static_dirty = true;
  std::vector<double> h_u, h_v, h_taux, h_tauy;
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0.0;
  int graph_n = 0;
  int graph_block = 0;
  int graph_lds = -1;
  int graph_prec = -1;
  std::vector<hipEvent_t> events;
  // measurement
  bool timing = false;
  double last_ms = 0.0, sum_ms = 0.0;
  int last_graph = 0;  // the last step replayed a captured hipGraph
  long long timed = 0, steps = 0;
  int block = 256;
  int lds_levels = 0;  // live levels of Thomas state kept in LDS per lane
  int lds_max = 65536;  // dynamic LDS a fused-step workgroup may take (the device's limit)
  long long live_levels = 0;  // sum over wet columns of live levels
  std::vector<int> col_live;  // live levels per grid column (from the last h_layer upload)
  bool block_pinned = false;
  int order_mode = 0;
  int precision = RDB_COLPHYS_PREC_FP64;
  int closure = RDB_COLPHYS_CLOSURE_PLAIN;
  std::vector<void*> allocations;
  // learned closure (trained or loaded network, dataset, guard state)
  double* d_coriolis = nullptr;
  colphys::SurrState* surr = nullptr;
  // momentum vertical viscosity: face velocities and fold inputs on the
  // device, created on first use
  colphys::MomState* mom = nullptr;
  long long h_gen = 0;  // bumped whenever h_layer is set (natural-order copies follow it)
  // the rdb_ocean.h host surface: simulation time, and host mirrors of the
  // fields the pointer getters hand out, each current as of state version
  // mirror_ver[name]; every entry point that changes the device state bumps
  // state_ver
  double t_sim = 0.0;
  long long state_ver = 0;
  std::map<std::string, std::vector<double>> mirrors;
  std::map<std::string, long long> mirror_ver;
  long long last_fallback = 0;
  double last_mlp_ms = 0.0;
  double mlp_flop = 0.0;  // per step, unpadded layer widths
  // batched ensemble: member plans (each its own grid and packed layout,
  // sharing this handle's stream), stepped together by one persistent launch
  std::vector<rdb_colphys*> members;
  bool own_stream = true;                  // members borrow the parent's stream
  colphys::StateView* d_views = nullptr;   // [members]
  colphys::SurrView* d_svs = nullptr;      // [members]
  int* d_nwet = nullptr;                   // [members]
  int2* d_work = nullptr;                  // (member, first packed column) per work item
  int n_work = 0;
  int work_block = 0;                      // block size the work list was cut for
  int persist_grid = 0;                    // workgroups of the persistent launch
  int nz_net = 0;                          // ensemble: the largest member nz (network levels)
  long long ens_wet = 0;                   // ensemble: wet columns over all members
  long long row_off = 0;                   // member: first row in the ensemble's network batch
  int lev_off = 0;                         // member: network levels below this grid's bed
  // contributors off at the defaults (EPBL, KV_ML_INVZ2, convection, double
  // diffusion): when any is on, every step runs the staged kernels
  bool extras = false;
  colphys::EpblParams ep{};
  colphys::EpblView ev{};                  // allocated when EPBL is on
  colphys::ContribParams cpar{};
};

namespace colphys {
// the member plans of a handle (itself when it is a single grid)
inline std::vector<rdb_colphys*> plans(rdb_colphys* h) {
  return h->members.empty() ? std::vector<rdb_colphys*>{h} : h->members;
}
int launch_step_all(rdb_colphys* h, double dt, bool surr);
int ensure_work(rdb_colphys* h);
}  // namespace colphys
