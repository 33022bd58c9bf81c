/*
 * minirender_amd.h -- C-ABI of the MI355X (gfx950) rasterization path.
 *
 * This is the boundary under minirender::Renderer::render(): plain C types,
 * plain pointers and sizes, an int status on every call (0 = ok, negative =
 * error), nothing thrown across it.  The reference is a C++ class with no FFI
 * (src/Renderer.cpp:311-381, include/minirender/Renderer.h:17-64); each entry
 * point below names the reference code whose work it takes over.
 *
 * Frame state that the reference derives on the host inside render()
 * (scene flattening Scene.cpp:13-36, modelview/normal matrix
 * Renderer.cpp:336-338, light to eye space :317-320) is computed by the
 * caller and passed in mr_frame / mr_draw, so the host math is shared by every
 * consumer of a frame.  The near-plane depth (:324-326) is derived here from
 * the projection, in fp32, as the reference does.
 */
#ifndef MINIRENDER_AMD_H
#define MINIRENDER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define MR_OK 0
#define MR_ERR_ARG -1      /* bad argument (null pointer, size out of range)   */
#define MR_ERR_HIP -2      /* a HIP runtime call failed                         */
#define MR_ERR_NOMEM -3    /* device allocation failed                          */
#define MR_ERR_BUDGET -4   /* request does not fit the context's HBM budget    */
#define MR_ERR_HANDLE -5   /* unknown mesh / texture / ticket handle            */
#define MR_ERR_NOGPU -6    /* no HIP device available                           */
#define MR_ERR_INDEX -7    /* mesh index out of range of its attribute array   */
#define MR_ERR_DEVICE -8   /* a kernel reported an internal error (spin limit) */
#define MR_ERR_TIMEOUT -9  /* a wait timed out                                   */
#define MR_ERR_CLOSED -10  /* server is shutting down                            */
#define MR_ERR_LIMIT -11   /* frame exceeds the fixed scratch of mr_render_noalloc */

/* ---- texture formats ----------------------------------------------------- */
#define MR_TEX_RGB8 0    /* 3 bytes per texel, texel value = byte / 255.0f (io.cpp:405-411) */
#define MR_TEX_RGB32F 1  /* 3 floats per texel, as Material::texture (Scene.h:42)         */

/* ---- one draw = one Renderable (Scene.h:47-53) with its material ------- */
/* mr_draw.flags */
#define MR_DRAW_NO_NEAR_CLIP 1 /* no near-plane test or clip: the triangles of
                                  Renderer::paintTriangle(a, b, c, world = false)
                                  (Renderer.cpp:169) */
typedef struct mr_draw {
  int32_t mesh;        /* handle from mr_mesh_upload                             */
  int32_t texture;     /* handle from mr_texture_upload, or -1 for none          */
  float modelview[16]; /* row-major, view * node transform (Renderer.cpp:337)     */
  float normalmat[16]; /* row-major, inverse(modelview) transposed (:338)         */
  float diffuse[3];    /* Material fields (Scene.h:38-45)                          */
  float specular[3];
  float emissive[3];
  float shininess;
  int32_t flags;       /* MR_DRAW_* (0: drawn as paintMesh draws, with clipping) */
} mr_draw;

/* ---- one frame = one Renderer::render() call --------------------------- */
typedef struct mr_frame {
  int32_t width, height;  /* Renderer::setSize (Renderer.cpp:100-106)           */
  float projection[16];   /* row-major, Renderer::setProjection                  */
  float background[3];    /* Renderer::setBackground, clear() (:113-119)         */
  float light[3];         /* eye-space light: view*light if point, else light   */
                          /* normalized (Renderer.cpp:317-320)                   */
  int32_t light_is_point; /* setLight(v, point)                                  */
  float ambient;          /* Scene::ambientLight (:322)                           */
  int32_t lighting;       /* setLighting                                          */
  int32_t texturing;      /* setTexturing                                         */
  int32_t save_normals;   /* setSaveNormals                                       */
  int32_t ndraws;
  const mr_draw* draws;   /* in Scene::collectShapes order (draw order)          */
} mr_frame;

/* Outputs of a frame.  Any pointer may be NULL (not produced).  With
 * on_device = 0 they are host pointers and the call copies results back;
 * with on_device = 1 they are device pointers on the context's device. */
typedef struct mr_outputs {
  uint8_t* rgb8;     /* H*W*3 RGB8, quantized as savePPM (io.cpp:358-361)         */
  float* depth;      /* H*W fp32 z-buffer, 1e11f where nothing was drawn (:116)  */
  uint32_t* winner;  /* H*W draw-order index of the visible triangle:              */
                     /* 2*(triangles of earlier draws + triangle) + clip sub-index */
                     /* (0/1); 0xFFFFFFFF where nothing was drawn                 */
  float* rgb;        /* H*W*3 unquantized colour, as Renderer::getImage (:383)    */
  float* normals;    /* H*W*3 eye-space normals, as getNormalsImage (h:63)        */
  float* points;     /* H*W*3 eye-space points from the depth buffer, as          */
                     /* Renderer::getRangeImage (Renderer.cpp:388-415)           */
  int32_t on_device;
} mr_outputs;

/* Per-frame statistics of the last call on a context. */
typedef struct mr_stats {
  int64_t triangles_in;       /* input triangles                               */
  int64_t triangles_clipped;  /* input triangles that crossed the near plane   */
  int64_t records;            /* triangles after clip/reject/cull (raster set) */
  int64_t bin_entries;        /* (tile, triangle) pairs                          */
  int64_t tiles;              /* screen tiles                                     */
  int64_t retries;            /* re-runs after a bin-capacity overflow           */
  double stage_ms[8];         /* per-stage GPU time when profiling is on:        */
                              /* 0 clear, 1 geometry, 2 scan, 3 bin, 4 rows      */
                              /* (edge checkpoints), 5 raster, 6 shade; for the   */
                              /* persistent kernel each stage's span (its first   */
                              /* workgroup's start to its last one's end)         */
  double frame_ms;            /* first to last kernel of the frame (profiling)  */
  int64_t launches;           /* kernel launches of the last batch (1: the     */
                              /* persistent kernel)                             */
  /* Captured batches (mr_ctx_set_graphs) over every slot of the context     */
  /* since its creation: replayed as captured, replayed after updating an    */
  /* instantiated graph of the same kernels in place, newly instantiated.    */
  int64_t graph_replays, graph_updates, graph_captures;
  int64_t split_tiles;        /* tiles split into parts of frag_size entries  */
  int64_t raster_parts;       /* parts and row bands rastered by workgroups    */
                              /* of their own                                   */
} mr_stats;

/* ---- device context: HIP device, stream, resident meshes/textures ------- */
typedef struct mr_ctx mr_ctx;

/* Number of HIP devices visible (0 without a GPU). */
int mr_device_count(int32_t* count);

/* Create a context on `device`.  hbm_budget bytes (0 = no limit) caps the
 * scratch and framebuffer memory the context may hold at once. */
int mr_ctx_create(int32_t device, uint64_t hbm_budget, mr_ctx** out);
int mr_ctx_destroy(mr_ctx* ctx);

/* Use an external HIP stream (hipStream_t) for all work of this context;
 * NULL restores the context's own stream. */
int mr_ctx_set_stream(mr_ctx* ctx, void* stream);
int mr_ctx_get_stream(mr_ctx* ctx, void** stream);

/* Tile configuration: tile_w x tile_h pixels per raster workgroup
 * (each of 8, 16, 32).  Defaults to 16 x 16. */
int mr_ctx_set_tile(mr_ctx* ctx, int32_t tile_w, int32_t tile_h);
/* Workgroup size of the tile raster for the current tile size (0 = the
 * default; 8x8: 64 or 128; 16x16 and 32x32: 64, 128, 256 or 512).
 * mr_ctx_set_tile resets it to the default. */
int mr_ctx_set_raster_threads(mr_ctx* ctx, int32_t threads);
/* Autotuner (no counterpart in the reference).  Renders `frames` (n >= 1,
 * with outputs `outs`) under every built combination of tile size (8, 16,
 * 32), raster workgroup size and split size of crowded tiles (256, 1024,
 * 4096 pairs), `reps` timed
 * runs each (HIP events on the context's stream), keeps the fastest and
 * sets the bin-list capacity to the measured pairs per triangle + 25%.  The
 * choice is applied to the context and returned in *best.  With cache_path,
 * a previous result for the same GPU and batch shape (frame count,
 * triangles, pixels, rounded up to powers of two) is read from that text
 * file instead of measuring, and new results are appended to it. */
typedef struct mr_tuning {
  int32_t tile;            /* tile edge in pixels                              */
  int32_t raster_threads;  /* raster workgroup size                            */
  float bin_factor;        /* bin-list entries per input triangle              */
  int32_t frag_size;       /* max (tile, triangle) pairs per raster work item  */
  double ms;               /* median time of one call with this choice         */
  int32_t from_cache;      /* 1 if read from cache_path                        */
  int32_t candidates;      /* configurations timed                             */
} mr_tuning;
int mr_ctx_autotune(mr_ctx* ctx, const mr_frame* frames, int32_t nframes,
                    const struct mr_outputs* outs, int32_t reps,
                    const char* cache_path, mr_tuning* best);
/* Largest number of (tile, triangle) pairs one raster workgroup takes; a
 * tile with more is split into parts merged through device memory
 * (64 .. 1<<20, default 1024). */
int mr_ctx_set_frag_size(mr_ctx* ctx, int32_t pairs);
/* Tiles whose walk work (bbox pixels of their entries inside the tile, plus 4
 * per row and 16 per entry) exceeds this are rastered as row bands, one
 * workgroup per band (0 = never; default 32768).  No counterpart in the
 * reference: load balance of the tile raster. */
int mr_ctx_set_band_work(mr_ctx* ctx, int32_t units);
/* Bin-list capacity in (tile, triangle) entries per input triangle
 * (>= 1); overflow re-runs the frame with a larger list. */
int mr_ctx_set_bin_factor(mr_ctx* ctx, float entries_per_triangle);
/* Capture the steady state as HIP graphs: a batch whose launch parameters
 * (sizes, scratch, tile configuration) equal an earlier one on the same
 * slot replays the captured uploads + kernels instead of launching them one
 * by one; a batch of other sizes with the same kernels updates such a graph
 * in place (hipGraphExecUpdate) and replays it, so a service of mixed
 * batches reuses a few instantiated graphs (mr_stats.graph_*).  Off by
 * default; ignored while profiling or on an external stream.
 * On MI355X with ROCm 7.2 a replayed single frame measured slower than eager
 * launches (which overlap the GPU's work); it is for callers whose host
 * thread is the bottleneck.  No counterpart in the reference (host-side
 * launch overhead). */
int mr_ctx_set_graphs(mr_ctx* ctx, int32_t on);

/* Run each batch of the synchronous calls (mr_render, mr_render_batch) as
 * ONE launch of a persistent kernel: a resident grid whose workgroups take
 * the stages' work from device work queues (one per XCD), each stage
 * starting once the stages before it are complete (bounded waits: a wait
 * over 2 s ends the batch with MR_ERR_DEVICE).  Same stages, bit-identical
 * results to the separate launches.  16 x 16 tiles only (MR_ERR_ARG
 * otherwise; a later tile change makes batches run as separate launches
 * again, see mr_stats.launches); the service and mr_render_noalloc always
 * use separate launches.  Off by default: measured on MI355X it is slower
 * at every size (1M-triangle 1080p frame 1.33 ms vs 0.148 ms; 64 such frames
 * 9.5 ms vs 4.6 ms), because each stage hand-over makes every workgroup
 * release and acquire at agent scope (an L2 write-back and invalidate per
 * workgroup on a GPU of eight L2s), where a kernel boundary does one of each.
 * Replaces the reference's per-frame call loop (samples/bench.cpp:128-135,
 * src/Renderer.cpp:311-332). */
int mr_ctx_set_persistent(mr_ctx* ctx, int32_t on);

/* How long one wait of the persistent kernel may take before the batch
 * gives up with MR_ERR_DEVICE, in microseconds (negative: the default,
 * 2 s).  0 makes every wait give up at once: the failure path can be
 * tested from the host while every kernel still exits normally.  No
 * counterpart in the reference. */
int mr_ctx_set_wait_limit(mr_ctx* ctx, int64_t us);
/* Record per-stage HIP event times into mr_stats (adds event overhead). */
int mr_ctx_set_profiling(mr_ctx* ctx, int32_t on);
int mr_ctx_get_stats(mr_ctx* ctx, mr_stats* stats);
/* Bytes of device memory currently held by the context. */
int mr_ctx_memory(mr_ctx* ctx, uint64_t* resident_bytes, uint64_t* scratch_bytes);

/* Upload a TriMesh (Scene.h:73-87): vertices, normals, texcoords and the three
 * index arrays (3 per triangle).  normal_indices may equal indices; uvs /
 * uv_indices may be NULL (then uv = (0,0) as in Renderer.cpp:371-379).
 * Indices are checked against the attribute counts (MR_ERR_INDEX).
 * Replaces the per-mesh copies of Renderer::paintMesh (:340-349). */
int mr_mesh_upload(mr_ctx* ctx, const float* positions, int32_t npositions,
                   const float* normals, int32_t nnormals, const float* uvs,
                   int32_t nuvs, const int32_t* indices,
                   const int32_t* normal_indices, const int32_t* uv_indices,
                   int32_t ntriangles, int32_t* handle);
/* Upload a TriMesh that uses the vertices, normals and texcoords of the
 * resident mesh `base`, as the per-material meshes of loadOBJ share one set
 * of attribute arrays (io.cpp:313-315).  Only the three index arrays are
 * uploaded (checked against base's attribute counts); uv_indices is ignored
 * when base has no texcoords.  Draws of meshes sharing attributes with the
 * same modelview in one frame share one set of vertex records, so the
 * vertex stage transforms each vertex once per frame instead of once per
 * mesh.  Releasing base keeps its attribute arrays until every mesh that
 * shares them is released. */
int mr_mesh_upload_shared(mr_ctx* ctx, int32_t base, const int32_t* indices,
                          const int32_t* normal_indices, const int32_t* uv_indices,
                          int32_t ntriangles, int32_t* handle);
int mr_mesh_release(mr_ctx* ctx, int32_t handle);

/* Upload a texture (Material::texture, Scene.h:42), rows x cols texels. */
int mr_texture_upload(mr_ctx* ctx, const void* texels, int32_t format,
                      int32_t rows, int32_t cols, int32_t* handle);
int mr_texture_release(mr_ctx* ctx, int32_t handle);

/* Render one frame: Renderer::render() (Renderer.cpp:311-332) for the draws
 * of `frame`.  Synchronous: returns when outputs are written. */
int mr_render(mr_ctx* ctx, const mr_frame* frame, const mr_outputs* out);

/* mr_render that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Outputs
 * must be device buffers (out->on_device = 1, else MR_ERR_ARG); meshes and
 * textures must already be resident.  All scratch of the frame comes from one
 * static device array of 64 MiB that each device the library runs on holds
 * until the process ends (a 1280x720 frame of ~300 k triangles fits; larger
 * frames get MR_ERR_LIMIT, a bin-list overflow MR_ERR_NOMEM instead of a
 * re-run with a larger list).  Every call on a device shares that array:
 * calls that overlap on one GPU (several threads or contexts) race.
 * Synchronous, on the context's stream; results equal mr_render's. */
int mr_render_noalloc(mr_ctx* ctx, const mr_frame* frame, const mr_outputs* out);

/* Render n independent frames in one batched pass (all stages launched once
 * for the whole batch).  Equivalent to n calls of mr_render. */
int mr_render_batch(mr_ctx* ctx, const mr_frame* frames, int32_t nframes,
                    const mr_outputs* outs);

/* ---- render service: continuous batching ------------------------------- */
/* No counterpart in the reference (it renders one frame per call,
 * Renderer.cpp:311): a request loop over the same pipeline.  Requests are
 * queued by mr_server_submit; a worker thread admits them in arrival order
 * into batches as execution slots and HBM budget free up, keeps up to
 * `slots` batches in flight on separate streams (so their work overlaps on
 * the device), and copies outputs back through pinned staging.  Each
 * request's result is identical to an mr_render of the same frame. */
typedef struct mr_server mr_server;

typedef struct mr_server_config {
  int32_t max_batch;     /* requests per batch (1 = plain per-request path)   */
  int32_t slots;         /* batches in flight at once (>= 1)                  */
  uint64_t hbm_budget;   /* bytes of framebuffer + scratch admitted at once   */
                         /* (0 = no limit)                                     */
  int32_t max_wait_us;   /* with a free slot, wait at most this long for more  */
                         /* requests before launching a partial batch (0)      */
} mr_server_config;

typedef struct mr_request_info {
  double submit_us;      /* submit time, microseconds since server start      */
  double launch_us;      /* its batch was launched                              */
  double done_us;        /* outputs written                                     */
  int64_t batch_id;
  int32_t batch_size;
  int32_t status;        /* MR_OK or MR_ERR_*                                   */
} mr_request_info;

int mr_server_create(mr_ctx* ctx, const mr_server_config* cfg, mr_server** out);
/* Stops accepting requests, finishes those queued, joins the worker. */
int mr_server_destroy(mr_server* srv);
/* Queue one frame.  The frame and its draws are copied; `out` must stay
 * valid until the request is done.  Meshes/textures must be resident in the
 * server's context.  Returns MR_ERR_BUDGET if the request alone exceeds the
 * budget. */
int mr_server_submit(mr_server* srv, const mr_frame* frame, const mr_outputs* out,
                     uint64_t* ticket);
/* done = 1 when the request's outputs are written. */
int mr_server_poll(mr_server* srv, uint64_t ticket, int32_t* done);
/* Wait for a request (timeout_ms < 0: no limit); fills info if non-NULL and
 * forgets the ticket.  Returns the request's status or MR_ERR_TIMEOUT. */
int mr_server_wait(mr_server* srv, uint64_t ticket, int32_t timeout_ms,
                   mr_request_info* info);
/* Replay an arrival trace: frame i is submitted at arrival_s[i] seconds
 * after the call starts; returns when all are done.  latency_ms[i] = done
 * time - scheduled arrival.  makespan_s = first arrival to last done. */
int mr_server_run_trace(mr_server* srv, const mr_frame* frames, int32_t n,
                        const double* arrival_s, const mr_outputs* outs,
                        double* latency_ms, double* makespan_s);

/* ---- host helpers of the reference API (no GPU needed) ----------------- */
/* Projection matrices, row-major, Renderer.cpp:40-83. */
int mr_projection_ortho(float l, float r, float b, float t, float n, float f,
                        float out[16]);
int mr_projection_perspective(float l, float r, float b, float t, float n,
                              float f, float out[16]);
int mr_projection_frustum(float fovy, float aspect, float n, float f,
                          float out[16]);
int mr_projection_frustum_h(float fovx, float aspect, float n, float f,
                            float out[16]);
int mr_projection_ortho_fov(float fov, float aspect, float n, float f,
                            float out[16]);
int mr_projection_cv(const float K[16], float w, float h, float n, float f,
                     float out[16]);
/* Near-plane eye depth derived from a projection (Renderer.cpp:324-326). */
int mr_znear(const float projection[16], float* znear);

#ifdef __cplusplus
}
#endif
#endif /* MINIRENDER_AMD_H */
