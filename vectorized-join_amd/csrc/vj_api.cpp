// C ABI entry points: status strings, hash kinds, vj_hash, and validation
// shared by the join entry points (vj_join_api.cpp).
#include <string.h>

#include <mutex>
#include <vector>

#include "vj_internal.hpp"

namespace vj {

static const char *kKindNames[HK_COUNT] = {
    "fnv", "custom_add", "custom_add_shift", "custom_add_hw", "custom_xor", "custom_xor_shift", "custom_xor_hw",
    "custom_mult", "custom_mult_shift", "custom_mult_hw", "custom_mult_xor_shift", "custom_n_mult_xor_shift",
    "custom_mult_xor_hw"};

int device_cu_count(int device) {
    static std::mutex mu;
    static std::vector<int> cache;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return 256;
    if ((int)cache.size() <= device) cache.resize(device + 1, 0);
    if (cache[device] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
            cus = 256;
        cache[device] = cus;
    }
    return cache[device];
}

Err check_table(const vj_table *t, bool) {
    if (!t) return Err::make(VJ_ERR_INVALID_ARGUMENT, "table is NULL");
    if (t->n_cols < 1 || t->n_cols > VJ_MAX_KEY_COLS) return Err::make(VJ_ERR_INVALID_ARGUMENT, "n_cols must be 1..4");
    if (t->key_width != 4 && t->key_width != 8) return Err::make(VJ_ERR_INVALID_ARGUMENT, "key_width must be 4 or 8");
    // Row, tile and chunk indexes are uint32 in the kernels: the end of the last
    // pass tile (6144 rows at most) and of the last probe chunk (16384 rows)
    // must stay below 2^32.
    if (t->n_rows > VJ_MAX_ROWS)
        return Err::make(VJ_ERR_UNSUPPORTED, "tables must have at most 2^32 - 2^16 rows");
    if (t->n_rows > 0)
        for (int c = 0; c < t->n_cols; c++)
            if (!t->cols[c]) return Err::make(VJ_ERR_INVALID_ARGUMENT, "key column pointer is NULL");
    return Err::ok();
}

Err make_hash_cfg(const vj_hash_opts *o, const vj_table *t, HashCfg *cfg) {
    vj_hash_opts d;
    vj_hash_opts_default(&d);
    if (!o) o = &d;
    if (o->kind < 0 || o->kind >= HK_COUNT) return Err::make(VJ_ERR_UNKNOWN_HASH, "unknown hash kind");
    if (o->hash_bits != 32 && o->hash_bits != 64) return Err::make(VJ_ERR_INVALID_ARGUMENT, "hash_bits must be 32 or 64");
    if (o->chunk_bits != 8 && o->chunk_bits != 32 && o->chunk_bits != 64 && o->chunk_bits != 128)
        return Err::make(VJ_ERR_INVALID_ARGUMENT, "chunk_bits must be 8, 32, 64 or 128");
    if (o->word_mode != VJ_WORDS_REFERENCE && o->word_mode != VJ_WORDS_NATIVE)
        return Err::make(VJ_ERR_INVALID_ARGUMENT, "unknown word_mode");
    cfg->kind = o->kind;
    cfg->bits = o->hash_bits;
    cfg->chunk_bits = o->chunk_bits;
    cfg->native = o->word_mode == VJ_WORDS_NATIVE;
    cfg->key_signed = t ? t->key_signed != 0 : 0;
    if (t) {
        // The row must be a whole number of chunks (join_provider.cuh:926-928).
        int words = cfg->native ? t->n_cols * t->key_width / 4 : t->n_cols * 2;
        if ((words * 32) % o->chunk_bits)
            return Err::make(VJ_ERR_INVALID_ARGUMENT,
                             "Columns do not fit fully into chunk buffer and padding is not provided");
    }
    return Err::ok();
}

TView to_view(const vj_table *t) {
    TView v;
    for (int c = 0; c < 4; c++) v.cols[c] = c < t->n_cols ? t->cols[c] : nullptr;
    v.rowids = t->rowids;
    v.n = t->n_rows;
    return v;
}

}  // namespace vj

using namespace vj;

extern "C" {

void vj_hash_opts_default(vj_hash_opts *o) {
    if (!o) return;
    o->kind = VJ_HASH_FNV;
    o->hash_bits = 64;
    o->chunk_bits = 64;
    o->word_mode = VJ_WORDS_REFERENCE;
}

const char *vj_status_str(vj_status s) {
    switch (s) {
    case VJ_OK: return "ok";
    case VJ_ERR_INVALID_ARGUMENT: return "invalid argument";
    case VJ_ERR_UNKNOWN_HASH: return "unknown hash algorithm";
    case VJ_ERR_UNSUPPORTED: return "unsupported";
    case VJ_ERR_BUDGET: return "HBM budget exceeded";
    case VJ_ERR_OUT_OF_MEMORY: return "out of device memory";
    case VJ_ERR_CAPACITY: return "output capacity too small";
    case VJ_ERR_HIP: return "HIP runtime error";
    case VJ_ERR_NO_DEVICE: return "no usable device";
    case VJ_ERR_NOT_FOUND: return "not found";
    case VJ_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

vj_status vj_hash_kind_from_name(const char *name, int32_t *kind) {
    if (!name || !kind) return VJ_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < HK_COUNT; k++) {
        if (strcmp(name, kKindNames[k]) == 0) {
            *kind = k;
            return VJ_OK;
        }
    }
    return VJ_ERR_UNKNOWN_HASH;
}

const char *vj_hash_kind_name(int32_t kind) {
    if (kind < 0 || kind >= HK_COUNT) return nullptr;
    return kKindNames[kind];
}

vj_status vj_hash(const vj_table *table, const vj_hash_opts *opts, void *out, void *stream) {
    Err e = check_table(table);
    if (e.failed()) return e.code;
    HashCfg cfg;
    e = make_hash_cfg(opts, table, &cfg);
    if (e.failed()) return e.code;
    if (table->n_rows == 0) return VJ_OK;
    if (!out) return VJ_ERR_INVALID_ARGUMENT;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return VJ_ERR_NO_DEVICE;
    e = launch_hash(table, cfg, out, (hipStream_t)stream, device_cu_count(dev));
    return e.code;
}

}  // extern "C"
