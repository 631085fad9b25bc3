// pz_exec.hpp -- batched gate execution over a device wire pool, shared by
// the FheUint8 circuit scheduler and the serving loop.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/pz.h"
#include "pz_internal.hpp"

namespace pz {

enum { GATE_NOT = 6 };

// One gate of a batch: op in 0..5 (nand..xnor, bootstrapped) or GATE_NOT
// (negation, no bootstrap); operands and result are rows of the wire pool;
// key indexes the batch's key list.
struct GateRec {
    u32 op, in0, in1, out, key;
};

// Runs every bootstrapped gate of `recs` in one prep / per-key key switch /
// persistent blind rotation sequence, then the NOT gates (in `recs` order,
// waves so that a NOT of a NOT sees its input).  Gates of one batch must not
// read wires written by the same batch's bootstrapped gates.
int exec_batch(pz_ctx *c, const std::vector<const pz_key *> &keys, std::vector<GateRec> &recs,
               u64 *pool, hipStream_t st);

}  // namespace pz

// context internals used by the executor (pz_capi.hip)
struct pz_exec_view {
    pz_params p;
    int logn;
    int8_t *digits;
    u32 *ksb, *ks, *stats;
    int grid;  // resident blind-rotation work groups (one-half kernel)
};
int pz_internal_exec_view(pz_ctx *c, size_t B, pz_exec_view *v, pz::DevCtx *dc);
// blind-rotation kernel variant (BR_ONE, BR_SPLIT_LAT, BR_SPLIT) and grid
// for a batch of B bootstraps
int pz_internal_br_plan(pz_ctx *c, size_t B, int *var, int *grid);
const pz::DevKey *pz_internal_devkey(const pz_key *k);
