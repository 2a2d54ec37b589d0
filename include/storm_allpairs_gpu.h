/* C++ mirror of the reference's batched all-pairs driver for the GPU engine.
 *
 * Same entry points, argument meaning and results as storm::allpairs_sum /
 * storm::allpairs_tiles (kernels/storm_allpairs.h:196-214), in namespace
 * storm::gpu, running on the MI355X engine through the C ABI (storm_hip.h).
 * Two forms:
 *   - this header, which does not depend on the reference's headers: rows as
 *     the reference's Row::list contents in CSR form (the engine rebuilds
 *     every representation with build_row's gates on upload), or an already
 *     uploaded corpus, on an explicit Engine;
 *   - storm_allpairs_gpu_ref.h, included after the reference's
 *     kernels/storm_allpairs.h: overloads with the reference's exact
 *     parameter lists (const std::vector<Row>&, const CostModel&, Policy, ...)
 *     on a process-wide default engine, uploading the Row arenas as built.
 *
 * allpairs_tiles streams: one tile row of the triangle is computed on the
 * device per step (storm_hip_allpairs_band) and handed to the callback tile
 * by tile, so neither side holds N^2 counts. Coordinates follow the
 * reference: density-order positions under PerTile (stable by cardinality,
 * kernels/storm_allpairs.cpp:316-323, with no permutation passed), row order
 * under every other policy.
 *
 * Policies: AllBitmap, PerTile (analytic per tile), Fixed, Probe (measure
 * then commit on the device), Refine (per-pair choice between a tile's two
 * best cells on heterogeneous tiles) and PerPair / Oracle (every pair picks
 * its own cell) -- storm_hip.h, STORM_HIP_POLICY_*.
 *
 * Differences a caller sees:
 *   - Probe measures tile throughput on sampled tiles and commits per class
 *     of tiles, not per tile on 3 timed pairs (a GPU's cost is a tile's);
 *   - Oracle is PerPair: the same model choice, made on the device;
 *   - the engine selects per tile of 64, 128 or 256 rows: a narrower `tile`
 *     (8, 16, 32) selects at 64 and only the callback granularity is `tile`;
 *   - a failure is reported in AllPairsStats::status (a storm_hip status
 *     code) with the message in storm_hip_last_error(); the reference has no
 *     failure path besides an empty result for fewer than two rows.
 */
#ifndef STORM_ALLPAIRS_GPU_H_
#define STORM_ALLPAIRS_GPU_H_

#include <cstddef>
#include <cstdint>
#include <vector>

#include "storm_hip.h"

namespace storm {
namespace gpu {

/* kernels/storm_allpairs.h:36-99 */
enum class Policy : uint8_t { AllBitmap, PerPair, PerTile, Oracle, Fixed, Probe, Refine };

/* Pairing order of kernels/storm_cost.h (BB..WW, Empty, CB). */
enum class Pairing : uint8_t { BB, BS, BR, BW, SS, SR, SW, RR, RW, WW, Empty, CB, COUNT };

const char* name_of(Policy p);

/* kernels/storm_allpairs.h:126-143, plus the status of the GPU call. */
struct AllPairsStats {
    uint64_t sum = 0;
    uint64_t pairs = 0;
    uint64_t decisions = 0;
    double ns_total = 0;
    double ns_selection = 0;
    uint64_t skipped = 0;
    uint64_t cell_pairs[(int)Pairing::COUNT] = {0};
    int status = STORM_HIP_OK;
};

/* Rows as sorted position lists: row r is pos[offs[r] .. offs[r+1]), every
 * position < universe. */
struct RowsCSR {
    const uint64_t* offs = nullptr;
    const uint32_t* pos = nullptr;
    uint32_t n_rows = 0;
    uint32_t universe = 0;
};

/* The engine a call runs on (one context, one device). */
class Engine {
public:
    explicit Engine(int device = 0);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    storm_hip_ctx* ctx() const { return ctx_; }
    int status() const { return status_; }

private:
    storm_hip_ctx* ctx_ = nullptr;
    int status_ = STORM_HIP_OK;
};

/* The engine of the reference-typed overloads: device 0, created on first use. */
Engine& default_engine();

/* Row arenas in the storm_repr layouts (storm_hip_arena), owned. Filled from
 * the reference's std::vector<Row> by storm_allpairs_gpu_ref.h. */
struct RowArenas {
    uint32_t n_rows = 0;
    uint32_t n_words = 0;
    std::vector<uint64_t> list_off, run_off, ewah_off, bitmap_off, rank_off, occ_off, comp_off;
    std::vector<uint32_t> list, run_start, run_end, comp, meta;
    std::vector<uint64_t> ewah, bitmap, rank, occ;
    bool has_ewah = true, has_comp = true;
};

/* A Row carries no universe (kernels/storm_repr.h:237-276). The smallest
 * universe U consistent with every row's arenas: ceil(U/64) = n_words, every
 * position < U, build_row's gates (zone map iff card*512 > U, complement iff
 * 2*card > U, storm_repr.cpp:139,188) and, where a complement exists,
 * U = card + |complement|. STORM_HIP_ERR_INVALID when none is. Drops the
 * EWAH / complement arenas when some row lacks them (sparse-only rows). */
int infer_universe(RowArenas& a, uint32_t* universe);

/* Upload arenas as given (storm_hip_corpus_upload). */
int upload_arenas(Engine& eng, RowArenas& a, storm_hip_corpus** out);

/* The device cost model a reference CostModel maps to: the engine's own
 * device-calibrated rates (the reference's are CPU ns), with every cell the
 * reference model leaves unpriced (ns_per_unit <= 0, excluded by
 * select_pairing, storm_cost.cpp:422-465) excluded on the device too. */
storm_hip_cost_model device_model_for(Engine& eng, const double* ns_per_unit, int n_cells);

/* kernels/storm_allpairs.h:196-201. `model` is the device cost model (null:
 * the context's current one); `no_zonemap` prices B x B at Theta(m) as the
 * device kernel runs it. */
AllPairsStats allpairs_sum(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile = 64, bool no_zonemap = false,
                           Pairing fixed_cell = Pairing::BS);

/* kernels/storm_allpairs.h:206-214: counts[(i - i0) * tile + (j - j0)] for the
 * tile's pairs i < j (density-order positions under PerTile, row order
 * otherwise), one call per tile of the upper triangle in the reference's
 * order, tile rows streamed from the device one at a time. */
typedef void (*TileCallback)(uint32_t i0, uint32_t j0, uint32_t nr, uint32_t nc,
                             const uint32_t* counts, void* ctx);

AllPairsStats allpairs_tiles(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile = 64);

/* The same two calls on a corpus already on the engine (any upload path). */
AllPairsStats allpairs_sum(Engine& eng, storm_hip_corpus* c, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile = 64, bool no_zonemap = false,
                           Pairing fixed_cell = Pairing::BS);
AllPairsStats allpairs_tiles(Engine& eng, storm_hip_corpus* c, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile = 64);

}  // namespace gpu
}  // namespace storm

#endif
