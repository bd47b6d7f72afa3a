// Host-side join execution shared by the context entry points (vj_context.cpp)
// and the serving engine (vj_server.cpp).
#pragma once
#include <stdint.h>

#include <vector>

#include "vj_internal.hpp"

// A partitioned build side kept in HBM: SoA key columns + row ids in
// partition order, and the 2^(b1+b2)+1 partition bases, in one allocation.
struct vj_built {
    int device = 0;
    uint64_t n = 0;
    int32_t n_cols = 0, key_width = 0, key_signed = 0;
    vj_hash_opts hash{};
    uint32_t b1 = 0, b2 = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    void *cols[4] = {};
    uint32_t *ids = nullptr;
    uint32_t *base = nullptr;
};

namespace vj {

// One join of a pipeline.
struct JoinSpec {
    const vj_table *r, *s;
    uint32_t *out_r, *out_s;
    uint64_t cap;
    const vj_built *res = nullptr;  // probe this resident build side (r unused)
    vj_built *keep = nullptr;       // keep r's partitioned rows here (allocated by the run)
};

// Device bytes a pipeline over `specs` needs: context scratch (with the
// arenas' growth slack) plus the tables it keeps.
Err pipeline_bytes(const std::vector<JoinSpec> &specs, const vj_join_opts *o, uint64_t *bytes);

// Launches the pipeline on `s` and the download of the match counts into the
// context's pinned buffer; nothing waits. `budget` caps the context's memory
// for this launch (0: none). The counts are read with pipeline_counts once
// the stream has passed this point.
Err pipeline_launch(vj_ctx *c, std::vector<JoinSpec> &specs, const vj_join_opts *o, hipStream_t s, uint64_t budget,
                    size_t *count_off);
void pipeline_counts(const vj_ctx *c, size_t count_off, size_t n, std::vector<uint64_t> &counts);

uint64_t ctx_bytes(const vj_ctx *c);
void ctx_release(vj_ctx *c);
void built_release(vj_built *b);

}  // namespace vj
