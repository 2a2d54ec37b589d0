/* C++ mirror of the reference's batched all-pairs driver for the GPU engine.
 *
 * Same entry points, argument meaning and results as storm::allpairs_sum /
 * storm::allpairs_tiles (kernels/storm_allpairs.h:196-214), in namespace
 * storm::gpu, running on the MI355X engine through the C ABI (storm_hip.h).
 * Rows are passed as the reference's Row::list contents in CSR form (the
 * engine rebuilds every representation with build_row's gates on upload), so
 * this header does not depend on the reference's headers.
 *
 * Differences a caller sees:
 *   - allpairs_tiles hands the callback tiles in ORIGINAL row order under
 *     every policy (the reference's PerTile hands density-sorted coordinates
 *     without the permutation, kernels/storm_allpairs.h:27-28);
 *   - policies PerPair and Oracle have no GPU form (status UNSUPPORTED);
 *     Probe and Refine map to the learned selector, which the caller trains
 *     first with storm_hip_selector_sample / storm_hip_selector_train on the
 *     same context (untrained, it falls back to the analytic model);
 *   - a failure is reported in AllPairsStats::status (a storm_hip status
 *     code) with the message in storm_hip_last_error(); the reference has no
 *     failure path besides an empty result for fewer than two rows.
 */
#ifndef STORM_ALLPAIRS_GPU_H_
#define STORM_ALLPAIRS_GPU_H_

#include <cstddef>
#include <cstdint>

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

/* kernels/storm_allpairs.h:196-201. `model` is the device cost model (null:
 * the context's current one); `no_zonemap` prices B x B at Theta(m) as the
 * device kernel runs it. */
AllPairsStats allpairs_sum(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile = 64, bool no_zonemap = false,
                           Pairing fixed_cell = Pairing::BS);

/* kernels/storm_allpairs.h:206-214: counts[(i - i0) * tile + (j - j0)] for the
 * tile's pairs i < j (original row order), one call per tile of the upper
 * triangle, tile rows streamed from the device one at a time. */
typedef void (*TileCallback)(uint32_t i0, uint32_t j0, uint32_t nr, uint32_t nc,
                             const uint32_t* counts, void* ctx);

AllPairsStats allpairs_tiles(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile = 64);

}  // namespace gpu
}  // namespace storm

#endif
