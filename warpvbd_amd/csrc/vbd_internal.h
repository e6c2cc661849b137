// vbd_internal.h -- solver state shared by the setup, step and C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/warpvbd.h"

struct vbd_mpnn;

namespace vbd {

constexpr int kMaxColors = 256;            // src/coloring.py:50
constexpr int kMaxElementsPerVertex = 32;  // src/vbd_solver.py:12
// neighbour staging of the big-colour sweep (256 threads per block)
constexpr int kStageRows = 64;    // vertices per block of 4-lane vertices (the batched sweep's rows)
// A staged block lists, vertex by vertex, each vertex's distinct neighbours
// (an entry per (vertex, neighbour) pair; 16 bytes of LDS each):
constexpr int kStageMax = 1152;      // entries of a block of 4-lane vertices (64 x degree 18)
constexpr int kStageMaxWide = 1408;  // entries of a block holding 2-lane vertices (128 per block)
constexpr int kStageEntryBits = 5;   // a vertex has at most 31 neighbours in its list (slot records' field)
constexpr int kStageDict = 64;    // dictionary entries staged with them (48 bytes each)
// A solver whose dictionary holds more entries (a batch of bodies of many
// materials: 20 entries per body) gives each staged block its own list of the
// at most kStageDict entries its slots use.
// Lanes per vertex of a single body's big colours, by slot count: at most
// kLaneSlots slots per lane.  A voxel mesh's interior vertices have 20 slots
// (4 lanes), its face vertices 10 (2 lanes) and its edge vertices 5 (1 lane),
// so every lane walks 5 slots and a colour needs fewer waves than at 4 lanes
// for every vertex.
constexpr int kLaneSlots = 5;
// The lane classes are used when at least this share of the big colours'
// vertices take 4 lanes (interior-dominated bodies: 88% on the 256x32x32
// beam).  Ordering vertices by class separates a body's surface vertices from
// its interior ones, which costs L2 locality; with many small bodies in one
// mesh (57% on 512 concatenated mass-springs) that cost outweighs the waves
// saved (measured 126 against 117 us per iteration), so such meshes keep 4
// lanes for every vertex.
constexpr double kLaneClassMinShare = 0.75;
// Colours below this many vertices are not staged: the staging pass adds a
// fixed latency per block, repaid only when a colour gives each of the 256
// CUs two or more blocks (measured per big-colour launch at 256 threads: 9.3k
// vertices 7.35 vs 6.31 us unstaged, 30k 9.14 vs 7.83, 47k 10.35 vs 10.67,
// 70k 13.05 vs 14.64).
constexpr int kStageMinColor = 40960;
// The tail sweep (k_sweep_tail): a single body's trailing colours of 32 lanes
// per vertex swept in one launch, one block (32 vertices of a colour x a few
// lanes each) per cluster of tail vertices joined by no edge to another
// cluster, each cluster's neighbours held in LDS.
constexpr int kTailNb = 1024;  // neighbours of a cluster (16 bytes of LDS each)
constexpr int kTailRows = 32;  // vertices of one colour in a cluster
constexpr int kMaxTail = 4;    // colours swept by one tail launch
// at most two blocks per CU: a colour of many vertices joined to no tail
// vertex would give a block to each of them (a 1024-thread block for a vertex)
constexpr int kTailMaxClusters = 512;

// Scratch of the allocation-free entry points (a file-scope device array in
// vbd_setup.hip, one copy per GPU for the life of the process).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

void set_error(const std::string &msg);
// a message with static storage, set without allocating
void set_error_literal(const char *msg);

#define VBD_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ::vbd::set_error(std::string(#call) + ": " + hipGetErrorString(e_));               \
            return VBD_EHIP;                                                                   \
        }                                                                                      \
    } while (0)

#define VBD_TRY(expr)                                                                          \
    do {                                                                                       \
        int r_ = (expr);                                                                       \
        if (r_ != 0) return r_;                                                                \
    } while (0)

// Per-step control block on the device.
struct Ctrl {
    int32_t it;         // index of the iteration being run
    int32_t converged;  // 1 once an iteration's metric fell below tol
    int32_t iters;      // iterations executed (index + 1 of the last one)
    uint32_t metric;    // bits of the last iteration's metric
    uint32_t blocks_done;  // blocks of the iteration's last launch that finished
    uint32_t launch_max;   // the tail sweep's running metric maximum over its blocks
};

// Per-body convergence state of the running step (batched solver).
struct BodyState {
    int32_t done;     // converged or out of iterations
    int32_t iters;    // iterations executed
    uint32_t metric;  // bits of the last iteration's max|dx_x+dx_y+dx_z|
    int32_t pad;
};

enum { kBatchLaunches = 1, kBatchResident = 2 };

struct CachedGraph {
    int k = 0, max_iter = 0, lanes = 0, block = 0;
    float tol = 0.f, kd_dt = 0.f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;  // graph_clock at the last replay
};
constexpr size_t kGraphCacheSize = 32;  // captured graphs a solver keeps (least recently used evicted)
constexpr int kMaxGraphIters = 32;       // iterations in the largest captured graph
// The largest power of two not above min(n, kMaxGraphIters) (n >= 1): a run of
// n iterations is replayed from graphs of 2^j iterations.
inline int graph_piece(int n)
{
    int p = 1;
    while (2 * p <= n && 2 * p <= kMaxGraphIters) p *= 2;
    return p;
}

struct Launch {
    int lanes = 4;        // lanes per vertex
    int block = 256;      // threads per block
    int small_color = 2048;  // colours below this many vertices get 32 lanes per vertex
    int chunk = 0;        // iterations between host waits (0 = adaptive)
    bool resident = true; // allow the one-workgroup whole-step kernel when every colour is small
    bool graph = true;    // batched colour launches: replay captured iterations
    bool compact = true;  // batched colour launches: after the first chunk, launch only live bodies' ranges
    int resident_block = 1024;  // threads per body in the one-workgroup-per-body form
};

// Lane class of a vertex with k slots in a single body's big colour: 0 = 4
// lanes (k > 2 kLaneSlots), 1 = 2 lanes, 2 = 1 lane (k <= kLaneSlots).  A
// colour's vertices are ordered by class (stable), so its lanes are
// [4 per class-0 vertex][2 per class-1 vertex][1 per class-2 vertex].
__host__ __device__ inline int lane_class(int k) { return k > 2 * kLaneSlots ? 0 : (k > kLaneSlots ? 1 : 2); }

// The launch layout of one big colour: n4, n2, n1 vertices of each class;
// 256-thread blocks over the colour's lanes; blocks below b4 hold only 4-lane
// vertices and stage up to kStageMax entries, blocks in [b4, b1) up to
// kStageMaxWide; blocks from b1 on (those holding 1-lane vertices, or every
// block when the colour has no staging lists) read their neighbours from
// global memory.  ids: first neighbour id of the colour's lists in st_nb.
struct LaneLayout {
    int n4 = 0, n2 = 0, n1 = 0;
    int b4 = 0, b1 = 0, nblocks = 0;
    int64_t ids = -1;  // -1: no staging lists
};

// Layout of a colour of nv vertices (n4, n2 of classes 0 and 1); staged: the
// lists exist (blocks below b1 use them).
inline LaneLayout lane_layout(int n4, int n2, int nv, bool staged, int64_t ids)
{
    LaneLayout l;
    l.n4 = n4;
    l.n2 = n2;
    l.n1 = nv - n4 - n2;
    const int64_t lanes = 4 * (int64_t)n4 + 2 * (int64_t)n2 + l.n1;
    l.nblocks = (int)((lanes + 255) / 256);
    // the first block holding a lane of a 2- or 1-lane vertex, then of a 1-lane one
    l.b4 = n2 + l.n1 > 0 ? (int)(4 * (int64_t)n4 / 256) : l.nblocks;
    l.b1 = !staged ? 0 : (l.n1 > 0 ? (int)((4 * (int64_t)n4 + 2 * (int64_t)n2) / 256) : l.nblocks);
    l.ids = staged ? ids : -1;
    return l;
}

// The colour-local vertex and its lane within the vertex's group for lane t
// of a layout (lg = log2 of the group's lanes).
__host__ __device__ inline int lane_vertex(int64_t t, int n4, int n2, int &g, int &lg)
{
    const int64_t t4 = 4 * (int64_t)n4, t2 = t4 + 2 * (int64_t)n2;
    if (t < t4) {
        g = (int)(t & 3);
        lg = 2;
        return (int)(t >> 2);
    }
    if (t < t2) {
        g = (int)((t - t4) & 1);
        lg = 1;
        return n4 + (int)((t - t4) >> 1);
    }
    g = 0;
    lg = 0;
    return n4 + n2 + (int)(t - t2);
}

// First entry of block b's list (blocks below b4 hold kStageMax entries, the
// rest kStageMaxWide).
__host__ __device__ inline int64_t stage_list_offset(int b, int b4)
{
    return b < b4 ? (int64_t)b * kStageMax : (int64_t)b4 * kStageMax + (int64_t)(b - b4) * kStageMaxWide;
}

// One colour's launch.
struct Segment {
    int c0, c1;    // colour range [c0, c1)
    int part_off;  // first per-block metric partial of this launch
    int nparts;    // blocks of this launch
    int lanes;     // lanes per vertex of this launch
    int block;     // threads per block of this launch
    int tail = 0;  // the tail sweep over colours [c0, c1) (k_sweep_tail), nparts clusters
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace vbd

struct vbd_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0, m = 0, material = 0;
    double damping = 0.0;
    double grav[3] = {0, 0, -9.81};

    // mesh facts
    int ncolors = 0, max_incident = 0, max_degree = 0, n_ref = 0;
    int64_t n_slots = 0;
    std::vector<int> col_start;   // colour-major start of each colour's active range (size ncolors+1)
    std::vector<int> col_size;    // active vertices per colour
    std::vector<double> col_bytes;  // algorithmic bytes of one sweep of each colour
    std::vector<double> col_slots, col_boundary;  // per colour: slots, distinct gathered neighbours
    std::vector<vbd::Segment> segments;
    int32_t *d_col_start = nullptr;  // device copy of col_start
    float kd_dt = 0.f;               // damping / dt of the running step
    bool resident = false;           // whole step in one workgroup (all colours small)
    int resident_lanes = 32;         // lanes per vertex of that form
    int n_bodies = 1;
    std::vector<int32_t> body_voff_host;  // body b owns original vertices [body_voff_host[b], body_voff_host[b+1])
    std::vector<int2> ranges_host;   // per body: (start, count) of each colour, then of its other vertices
    int2 *d_ranges = nullptr;        // device copy of ranges_host
    int32_t *body_new = nullptr;     // body of each vertex, colour-major order
    bool batched = false;            // built by vbd_batch_create
    int batch_mode = vbd::kBatchLaunches;
    vbd::BodyState *bstate = nullptr, *bstate_host = nullptr;
    uint32_t *dxm = nullptr;         // per body: metric bits of the current iteration (max over its vertices)
    int32_t *n_live = nullptr, *n_live_host = nullptr;
    // Rows of kStageRows consecutive vertices of a colour, numbered colour by
    // colour from row_base[c]: per row the number of live bodies with a
    // vertex in it.  A launch skips the rows whose count is zero, so a body's
    // convergence removes its rows from the remaining launches of the step
    // without the host building lists.
    std::vector<int32_t> row_base;     // per colour: first row (size ncolors+1)
    int n_rows = 0;
    int32_t *row_live = nullptr;       // [n_rows] during a step
    int32_t *row_live0 = nullptr;      // [n_rows] at the start of a step (every body live)
    int32_t *body_row_off = nullptr;   // [n_bodies+1] per body: its rows, in body_rows
    int32_t *body_rows = nullptr;
    // live-range tasks of the batched colour launches: per colour, blocks of
    // (first vertex, count) covering the bodies live when the list was built
    std::vector<int2> tasks_host;
    std::vector<int> task_off;       // per colour: first task, then the total
    int2 *d_tasks = nullptr;
    size_t task_cap = 0;
    int64_t task_vertices = 0;       // vertices the current list covers
    int median_iters = 0;            // median body iteration count of the last step
    std::vector<vbd::CachedGraph> graphs;
    uint64_t graph_clock = 0;
    int32_t *tets_dev = nullptr;     // elements (original ids), kept for the warm-start graph
    double edge_length = 0.01;       // cube root of 6 x mean tet volume
    vbd_mpnn *warm = nullptr;        // warm-start network used by the running step
    // tgt still holds the inertia target of the step that produced uprev
    // (set by every finished step, cleared by create, reset and snapshot
    // loads), so uprev - tgt is that step's converged correction
    bool hist_valid = false;
    double hist_dt = 0.0;            // dt of that step
    // serial number of the last finished step; reset and snapshot loads skip
    // one, so that no step before them is taken as the one right before the next
    int64_t step_serial = 0;
    // changes made from outside a step: positions or velocities set (or
    // passed to a step), material, gravity, reset, snapshot loads
    uint64_t edits = 0;
    double uniform_E = 0.0, uniform_nu = 0.0;  // the material when every element shares it (else NaN)
    float4 *d_grav = nullptr;        // per body gravity
    int2 *d_body_out = nullptr;      // per body (iterations, metric bits) of the last step
    int2 *body_out_host = nullptr;   // pinned
    int n_swept = 0;              // active, coloured vertices (prefix of the colour-major order)

    // device arrays (colour-major order unless noted)
    int32_t *perm = nullptr;      // new -> original
    int32_t *iperm = nullptr;     // original -> new
    double *X = nullptr;          // rest positions, original order, f64 [n*3]
    // Vertex state beyond FP32 (the per-slot arithmetic is FP32): an FP32
    // displacement has a 1e-9 m ulp at 1 cm, which moves a tolerance-
    // terminated trajectory off the FP64 one by iteration-count flips.  The
    // iterate u is packed to 16 bytes (one load per neighbour gather); the
    // step's start and target, read once per vertex, are FP64.
    float4 *u = nullptr;          // displacement x - X, packed (32 significant bits, pack_disp in vbd_sweep.h)
    double4 *uprev = nullptr;     // displacement at the start of the step
    float4 *vel = nullptr;        // velocity
    double4 *tgt = nullptr;       // xyz: inertia target (x_prev + dt v + dt^2 g) - X ; w: M_v / dt^2
    float *mass = nullptr;        // lumped vertex mass sum_e m_e / 4
    uint8_t *active = nullptr;    // active flag
    int32_t *slot_off = nullptr;  // [n_swept+1]
    int32_t *slot_ent = nullptr;  // [n_slots] tet << 2 | local index (tet-record layout)
    float4 *slots = nullptr;      // [n_slots] compact slot records (ids, entry), or tet records [m*4]
    float2 *slots2 = nullptr;     // full records' qd while they are built; null afterwards
    float4 *dict = nullptr;       // [n_dict*3] record dictionary (vbd_device.h)
    float4 *dict_ext = nullptr;   // [n_dict*3] rest-edge constants of each entry (vbd_device.h)
    // Per colour: the vertices of lane classes 0 and 1 (lane_class), which the
    // colour-major order puts first (single-body solvers; a batch keeps one
    // class, (col_size, 0)).
    std::vector<int2> col_n42;
    // Neighbour staging for the 256-thread sweep of colours of at least
    // kStageMinColor vertices (shared dictionary): per block one entry per
    // (vertex, neighbour), per vertex its first entry, and per slot a 4-byte
    // record of indices into the vertex's entries plus the dictionary entry.
    bool st_ok = false;
    std::vector<vbd::LaneLayout> st_col;  // per colour: its layout and lists (nblocks 0: not a big colour)
    std::vector<int64_t> st_ids;   // per colour: entries its blocks stage (sum over blocks)
    int32_t *st_nb = nullptr;      // entry lists (stage_list_offset): neighbour id | owner << 25, -1 past a list
    // the distinct slot-record sequences of the staged vertices, one after
    // another, 32 zero records past the last: per slot n0 | n1 << 5 | n2 << 10
    // | entry << 15 (n_k: indices into the vertex's list)
    uint32_t *st_rec = nullptr;
    // [n_swept] per staged vertex: its first entry in its block's list | its
    // slot count << 16, and its record sequence's first record in st_rec
    int2 *st_vrec = nullptr;
    std::vector<double> st_prec;   // per colour: records of the distinct sequences its vertices use
    int32_t *st_dl = nullptr;      // per staged block kStageDict dictionary entries, -1 padded (null: shared)
    int32_t n_dict = 0;
    bool dict_shared = false;     // entries shared between slots (else one per slot)
    // Colours swept with 32 lanes per vertex (below launch.small_color): their
    // compact records padded to 32 per vertex (entry -1 past the vertex's
    // slots), so a lane finds its record without the slot offset.
    float4 *small_rec = nullptr;
    std::vector<int64_t> small_off;  // per colour: first row of 32 records, or -1
    // The tail sweep (single body): the colours tail_col[0..tail_k) swept by
    // tail_clusters blocks.  Per (cluster, colour k, row) the row's vertex and
    // that vertex's index in the cluster's neighbour table (-1: none); per row
    // and lane the lane's slot as three table indices and its dictionary
    // entry (-1 past the vertex's slots); per cluster its neighbours' ids,
    // tail_nb_stride per cluster, -1 past its count.
    bool tail_ok = false;
    int tail_k = 0, tail_clusters = 0, tail_nb_stride = 0;
    int tail_col[vbd::kMaxTail] = {};
    int tail_rows[vbd::kMaxTail] = {};  // rows of colour k per cluster (the most any cluster uses)
    int2 *tail_v = nullptr;             // [cluster][sum of tail_rows]
    uint2 *tail_rec = nullptr;          // [cluster][sum of tail_rows][32]: n0 | n1 << 10 | n2 << 20, entry
    int32_t *tail_nb = nullptr;         // [cluster][tail_nb_stride]
    int32_t *colors = nullptr;    // original order
    int32_t *inc_off = nullptr;   // original order vertex -> incident (tet<<2|local), ascending tet
    int32_t *inc = nullptr;
    uint32_t *part = nullptr;     // per-block metric maxima of one iteration
    int part_cap = 0, n_parts = 0;
    vbd::Ctrl *ctrl = nullptr;
    vbd::Ctrl *ctrl_host = nullptr;  // pinned
    double *xio = nullptr;        // f64 [n*3] staging for host transfers (device)
    double *xio_host = nullptr;   // pinned f64 [n*3]
    double4 *snap[VBD_MAX_SNAPSHOTS] = {};

    vbd::Launch launch;
    // a step that runs its last allowed iteration keeps the iterate from
    // before it (u_last), and the largest |dx| component of that iteration is
    // reported (last_dx_max, NaN when the last step stopped earlier)
    float4 *u_last = nullptr;
    double *d_dx_max = nullptr;
    double last_dx_max = 0.0;
    int last_iters = 0;
    int prev_iters = 0;  // iterations of the step before the last (single-body launches)

    // profiling
    bool prof = false;
    int prof_min = 0;
    std::vector<hipEvent_t> ev_pool;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against a caller's stream (vbd_step_dev)
    std::vector<std::pair<int, int>> ev_pending;  // (event index, colour)
    double prof_ms = 0.0, prof_bytes = 0.0;
    int64_t prof_launches = 0;
};

namespace vbd {
// The graph cache: a replayed entry is marked used; before a capture, the
// least recently used entry goes once kGraphCacheSize are held.
inline void graph_touch(vbd_solver *s, CachedGraph &g) { g.used = ++s->graph_clock; }
inline void graph_make_room(vbd_solver *s)
{
    if (s->graphs.size() < kGraphCacheSize) return;
    auto old = s->graphs.begin();
    for (auto it = s->graphs.begin(); it != s->graphs.end(); ++it)
        if (it->used < old->used) old = it;
    (void)hipGraphExecDestroy(old->exec);
    (void)hipGraphDestroy(old->graph);
    s->graphs.erase(old);
}
int setup_mesh(vbd_solver *s, const double *rest_xyz, const int32_t *tets, const double *density,
               const double *youngs, const double *poisson, const uint8_t *active, const int32_t *body_voff,
               int n_bodies, const double *body_grav);
int color_tets_device(const int32_t *d_tets, int m, int n, int32_t *d_colors, int32_t *d_inc_off,
                      int32_t *d_inc, int *ncolors, int *max_degree, hipStream_t stream);
// color_tets_device on caller-owned device buffers, every intermediate array
// in the scratch array (no allocation); kernels on `stream`
int color_tets_noalloc(const int32_t *d_tets, int m, int n, int32_t *d_colors, int *ncolors, int *max_degree,
                       hipStream_t stream);
int color_graph_device(const int32_t *d_off, const int32_t *d_adj, int n, int32_t *d_colors, int *ncolors,
                       hipStream_t stream);
int step_device(vbd_solver *s, double dt, double tol, int max_iter, int *iters, double *metric);
int export_positions(vbd_solver *s, const double4 *src, double *host_out);
int export_positions(vbd_solver *s, const float4 *src, double *host_out);  // packed displacements
int import_positions(vbd_solver *s, const double *host_in, double4 *dst);
int import_velocities(vbd_solver *s, const double *host_in);
// device f64 [n*3] buffers in original vertex order, enqueued on s->stream
int export_positions_dev(vbd_solver *s, const double4 *src, double *dev_out);
int export_positions_dev(vbd_solver *s, const float4 *src, double *dev_out);
int import_positions_dev(vbd_solver *s, const double *dev_in, double4 *dst);
int export_velocities(vbd_solver *s, double *host_out);
void plan_segments(vbd_solver *s);
// padded records of the colours plan_segments gives 32 lanes, then the tail
// sweep's tables (and the plan again, with the tail as one launch)
int build_small_records(vbd_solver *s);
int rescale_material(vbd_solver *s, double E0, double nu0, double E1, double nu1);
int build_staging(vbd_solver *s);  // neighbour staging tables (st_*), when the mesh allows
// The lane layout of colour c's launch when it runs the big-colour sweep
// (single-body solver, 4 lanes, 256 threads, at least kStageMinColor
// vertices), else nullptr.
const LaneLayout *big_color_layout(const vbd_solver *s, int c);
extern bool g_record_dictionary;  // share record dictionary entries in later creates (default on)
extern bool g_staging;            // neighbour staging in later creates (default on)
extern bool g_tail;               // the tail sweep in later creates and launch changes (default on)
extern bool g_edge_form;          // rest-edge constants for staged dictionaries in later creates (default on)
hipError_t rocprim_exclusive_scan_i32(void *tmp, size_t &bytes, const int32_t *in, int32_t *out, int n,
                                      hipStream_t st);
int step_bodies(vbd_solver *s, double dt, double tol, int max_iter);
void release_batch(vbd_solver *s);
int autotune_bodies(vbd_solver *s, double dt, double tol, int max_iter, int trial_steps, double *best_ms);
}  // namespace vbd
