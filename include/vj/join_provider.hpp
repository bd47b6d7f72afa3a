/*
 * vj/join_provider.hpp -- the reference's JoinProvider API over libvj.so.
 *
 * Callers written against vgerber/vectorized-join's join path
 * (JoinProvider, JoinConfig, JoinStatus, JoinSummary and db_table from
 * src/join/join_provider.cuh:54-239 and src/base/types.hpp:128-266, used as
 * in src/join_benchmark.cu:209-275) compile against this header unchanged and
 * run on the MI355X library through its C ABI (vj.h). Only the names,
 * fields and meaning of that API are kept; everything behind it is libvj.
 *
 * What each piece maps to:
 *  - db_table: the reference's row-major table (64-bit key columns
 *    column_values[row * column_count + c], 64-bit primary_keys). Device
 *    tables are allocated with hipMalloc, host tables with new[].
 *  - JoinProvider::join(r, s, rs): vj_join_rows. rs receives column_count 2,
 *    (r pk, s pk) per matching pair and primary keys 1..M, allocated by the
 *    callee; the caller frees it with rs.free(), as in the reference.
 *    Host tables (gpu == false) are copied to the device, joined there and
 *    the result copied back (the reference runs a brute-force CPU join,
 *    join_provider.cuh:241-284; the pair set is the same).
 *  - JoinConfig: `buckets` sets the radix bits per pass (ceil(log2(buckets)),
 *    join_provider.cuh:296-297); hash_config.algorithm picks the hash by the
 *    reference's names; HASH_BITS / HASH_CHUNK_BITS (src/base/config.h:28,31)
 *    are honoured when the caller defines them (default 64 / 64).
 *    devices must be 1 (one GPU). tasks_p_device, vector_bytes_size,
 *    join_mode, partition_config and probe_config are accepted and unused:
 *    the library sizes its own launches, LDS tables and work queue.
 *  - JoinSummary: tuples/s per phase from the library's per-stage device
 *    times (hashing is fused into the histogram and pass kernels, so the hash
 *    rate is the histogram stage's); join rates from the wall clock around
 *    the call, with the reference's join GB/s formula
 *    (join_provider.cuh:409-414).
 *
 * Errors are JoinStatus{false, message}, never exit().
 */
#ifndef VJ_JOIN_PROVIDER_HPP
#define VJ_JOIN_PROVIDER_HPP

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../vj.h"

typedef unsigned long long index_t;
typedef unsigned long long column_t;

struct JoinStatus {
    std::string message = "Unknown";
    bool successful = false;

    JoinStatus() = default;
    JoinStatus(bool ok, std::string msg) : message(std::move(msg)), successful(ok) {}
    explicit JoinStatus(bool ok) : message(ok ? "Successful" : "Failed"), successful(ok) {}

    bool is_successful() const { return successful; }
    bool has_failed() const { return !successful; }
};

struct db_table {
    size_t column_count = 0;
    column_t *primary_keys = nullptr;
    column_t *column_values = nullptr;
    index_t size = 0;
    bool gpu = false;
    bool data_owner = true;
    int device_index = 0;

    db_table() = default;

    // An owning table of `table_size` rows on the device (gpu) or the host.
    db_table(size_t columns, index_t table_size, bool on_gpu = true)
        : column_count(columns), size(table_size), gpu(on_gpu), data_owner(true) {
        if (gpu) {
            if (hipMalloc(&column_values, bytes_of(columns * table_size)) != hipSuccess ||
                hipMalloc(&primary_keys, bytes_of(table_size)) != hipSuccess) {
                std::cerr << "db_table: hipMalloc failed" << std::endl;
                column_values = primary_keys = nullptr;
                size = 0;
            }
        } else {
            column_values = new column_t[columns * table_size];
            primary_keys = new column_t[table_size];
        }
    }

    // A non-owning view of rows [offset, offset + rows) of `source`.
    db_table(index_t offset, index_t rows, const db_table &source)
        : column_count(source.column_count), size(rows), gpu(source.gpu), data_owner(false),
          device_index(source.device_index) {
        column_values = source.column_values + offset * column_count;
        primary_keys = source.primary_keys + offset;
    }

    // An owning copy on the device (to_gpu) or the host.
    db_table copy(bool to_gpu) const {
        db_table t(column_count, size, to_gpu);
        t.device_index = device_index;
        const hipMemcpyKind kind = gpu ? (to_gpu ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost)
                                       : (to_gpu ? hipMemcpyHostToDevice : hipMemcpyHostToHost);
        if (size) {
            (void)hipMemcpy(t.column_values, column_values, bytes_of(column_count * size), kind);
            (void)hipMemcpy(t.primary_keys, primary_keys, bytes_of(size), kind);
        }
        return t;
    }

    size_t get_bytes() const { return get_bytes(size, (int)column_count); }
    static size_t get_bytes(index_t elements, int columns) { return sizeof(column_t) * elements * (columns + 1); }

    void free() { free(nullptr); }
    void free(hipStream_t stream) {
        if (size > 0 && data_owner) {
            if (gpu) {
                if (stream) (void)hipStreamSynchronize(stream);
                (void)hipFree(column_values);
                (void)hipFree(primary_keys);
            } else {
                delete[] column_values;
                delete[] primary_keys;
            }
        }
        column_count = 0;
        size = 0;
        column_values = primary_keys = nullptr;
    }

    void print() const {
        db_table h = gpu ? copy(false) : *this;
        for (index_t i = 0; i < h.size; i++) {
            std::cout << i << "\t\t" << h.primary_keys[i] << "\t\t";
            for (size_t c = 0; c < h.column_count; c++) std::cout << h.column_values[i * column_count + c] << "\t\t";
            std::cout << std::endl;
        }
        if (gpu) h.free();
    }

  private:
    static size_t bytes_of(index_t n) { return (size_t)n * sizeof(column_t); }
};

struct HashSummary {
    std::string algorithm = "fnv";
    index_t elements = 0;
    int element_bytes = 0;
    int element_offset = 0;
    float k_tuples_p_seconds = 0.0f;
    float k_gb_p_seconds = 0.0f;
};

struct HashConfig {
    int threads_per_block = 256;  // accepted, unused: the library sizes its launches
    int elements_per_thread = 1;  // accepted, unused
    std::string algorithm = "fnv";
    hipStream_t stream = nullptr;
    bool profile_enabled = false;
    HashSummary profile_summary;
};

struct PartitionSummary {
    int device_index = 0;
    index_t elements = 0;
    float k_histogram_elements_p_second = 0.0f;
    float k_swap_elements_p_second = 0.0f;
    std::string to_string() const {
        return std::to_string(device_index) + "," + std::to_string(elements) + "," +
               std::to_string(k_histogram_elements_p_second) + "," + std::to_string(k_swap_elements_p_second);
    }
};

struct PartitionConfig {};  // accepted, unused: the planner chooses passes and bits

struct ProbeSummary {
    index_t r_elements = 0, s_elements = 0, rs_elements = 0;
    float k_probe_tuples_p_second = 0.0f;
};

struct ProbeConfig {  // accepted, unused: build/probe run in LDS tables sized by the library
    int probe_mode = 0;
    int build_n_per_thread = 1;
    float build_table_load = 0.25f;
    int build_threads = 256;
    int extract_n_per_thread = 1;
    int extract_threads = 256;
    void *d_probe_buffer = nullptr;
};

struct JoinSummary {
    JoinStatus join_status;

    float hash_tuples_p_second = 0.0f;
    std::vector<HashSummary> hash_summaries;
    float partition_tuples_p_second = 0.0f;
    std::vector<PartitionSummary> partition_summaries;
    float probe_tuples_p_second = 0.0f;
    std::vector<ProbeSummary> probe_summaries;
    float merge_tuples_p_second = 0.0f;
    float merge_gb_p_second = 0.0f;
    float join_tuples_p_second = 0.0f;
    float join_gb_p_second = 0.0f;

    index_t r_elements = 0;
    index_t s_elements = 0;
    index_t rs_elements = 0;

    JoinSummary &operator+=(const JoinSummary &js) {
        hash_tuples_p_second += js.hash_tuples_p_second;
        partition_tuples_p_second += js.partition_tuples_p_second;
        probe_tuples_p_second += js.probe_tuples_p_second;
        merge_tuples_p_second += js.merge_tuples_p_second;
        join_tuples_p_second += js.join_tuples_p_second;
        join_gb_p_second += js.join_gb_p_second;
        rs_elements += js.rs_elements;
        join_status = js.join_status;
        return *this;
    }
    JoinSummary &operator/=(const float &factor) {
        hash_tuples_p_second /= factor;
        partition_tuples_p_second /= factor;
        probe_tuples_p_second /= factor;
        merge_tuples_p_second /= factor;
        join_tuples_p_second /= factor;
        join_gb_p_second /= factor;
        rs_elements = (index_t)(rs_elements / factor);
        return *this;
    }
    void print() const {
        std::cout << "Hash      " << hash_tuples_p_second << " Tuples/s" << std::endl;
        std::cout << "Partition " << partition_tuples_p_second << " Tuples/s" << std::endl;
        std::cout << "Probe     " << probe_tuples_p_second << " Tuples/s" << std::endl;
        std::cout << "Merge     " << merge_tuples_p_second << " Tuples/s" << std::endl;
        std::cout << rs_elements << " Joins " << join_tuples_p_second << " Tuples/s" << join_gb_p_second << " GB/s"
                  << std::endl;
    }
};

struct JoinConfig {
    static const int MODE_RADIX_SPLIT = 0;
    static const int MODE_VECTOR_SPLIT = 1;

    int tasks_p_device = 1;
    int devices = 1;
    bool profile_enabled = true;
    int vector_bytes_size = 0;
    int join_mode = MODE_RADIX_SPLIT;
    int buckets = 2;

    HashConfig hash_config;
    PartitionConfig partition_config;
    ProbeConfig probe_config;
};

class JoinProvider {
  public:
    explicit JoinProvider(JoinConfig config) : join_config(config) {
        const vj_status st = vj_ctx_create(0, &ctx);
        if (st != VJ_OK) {
            ctx = nullptr;
            create_error = std::string("vj_ctx_create: ") + vj_status_str(st);
        } else {
            (void)vj_ctx_set_profiling(ctx, join_config.profile_enabled ? 1 : 0);
        }
    }
    ~JoinProvider() {
        if (ctx) (void)vj_ctx_destroy(ctx);
    }
    JoinProvider(const JoinProvider &) = delete;
    JoinProvider &operator=(const JoinProvider &) = delete;

    JoinStatus join(db_table r_table, db_table s_table, db_table &joined_rs_table) {
        join_summary = JoinSummary();
        if (r_table.gpu != s_table.gpu) return finish(JoinStatus(false, "Unable to figure out the join algorithm"));
        if (r_table.column_count != s_table.column_count)
            return finish(JoinStatus(false, "r and s tables have different column counts"));
        if (!ctx) return finish(JoinStatus(false, create_error));
        if (join_config.devices != 1) return finish(JoinStatus(false, "this build joins on one GPU (devices = 1)"));
        if (!r_table.gpu) {  // host tables: join on the device, result back on the host
            db_table r = r_table.copy(true), s = s_table.copy(true), rs;
            const JoinStatus st = join(r, s, rs);
            r.free();
            s.free();
            if (st.has_failed()) return st;
            joined_rs_table = rs.copy(false);
            rs.free();
            return st;
        }
        vj_join_opts opts;
        vj_join_opts_default(&opts);
        int32_t kind = 0;
        if (vj_hash_kind_from_name(join_config.hash_config.algorithm.c_str(), &kind) != VJ_OK)
            return finish(JoinStatus(false, "unknown hash algorithm " + join_config.hash_config.algorithm));
        opts.hash.kind = kind;
#ifdef HASH_BITS
        opts.hash.hash_bits = HASH_BITS;
#endif
#ifdef HASH_CHUNK_BITS
        opts.hash.chunk_bits = HASH_CHUNK_BITS;
#endif
        if (join_config.buckets > 1) {
            const int radix_width = (int)std::ceil(std::log2((double)join_config.buckets));
            opts.bits1 = opts.bits2 = radix_width < 1 ? 1 : (radix_width > 8 ? 8 : radix_width);
        }
        vj_rows r{r_table.size, (int32_t)r_table.column_count, 0, as_u64(r_table.column_values),
                  as_u64(r_table.primary_keys)};
        vj_rows s{s_table.size, (int32_t)s_table.column_count, 0, as_u64(s_table.column_values),
                  as_u64(s_table.primary_keys)};
        vj_rows rs{};
        const auto t0 = std::chrono::high_resolution_clock::now();
        const vj_status st = vj_join_rows(ctx, &r, &s, &opts, &rs, join_config.hash_config.stream);
        const double wall = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        if (st != VJ_OK) return finish(JoinStatus(false, std::string(vj_status_str(st)) + ": " + vj_last_error(ctx)));
        joined_rs_table = db_table();
        joined_rs_table.column_count = 2;
        joined_rs_table.size = rs.n_rows;
        joined_rs_table.column_values = reinterpret_cast<column_t *>(rs.column_values);
        joined_rs_table.primary_keys = reinterpret_cast<column_t *>(rs.primary_keys);
        joined_rs_table.gpu = true;
        joined_rs_table.data_owner = true;
        fill_summary(r_table, s_table, rs.n_rows, wall);
        return finish(JoinStatus(true));
    }

    JoinSummary get_join_summary() const { return join_summary; }

  private:
    // column_t (unsigned long long) and uint64_t are both 64-bit unsigned
    static uint64_t *as_u64(column_t *p) { return reinterpret_cast<uint64_t *>(p); }

    JoinStatus finish(JoinStatus st) {
        join_summary.join_status = st;
        return st;
    }

    void fill_summary(const db_table &r, const db_table &s, index_t m, double wall_s) {
        vj_join_summary vs;
        std::memset(&vs, 0, sizeof(vs));
        (void)vj_join_summary_get(ctx, &vs);
        const double n = (double)(r.size + s.size);
        auto rate = [n](double ms) { return ms > 0 ? (float)(n / (ms * 1e-3)) : 0.0f; };
        join_summary.hash_tuples_p_second = rate(vs.ms_histogram);
        join_summary.partition_tuples_p_second = rate(vs.ms_histogram + vs.ms_pass1 + vs.ms_pass2);
        join_summary.probe_tuples_p_second = rate(vs.ms_build_probe);
        join_summary.join_tuples_p_second = wall_s > 0 ? (float)(n / wall_s) : 0.0f;
        join_summary.join_gb_p_second =
            wall_s > 0 ? (float)(n * (double)r.column_count * sizeof(column_t) / 1e9 / wall_s) : 0.0f;
        join_summary.r_elements = r.size;
        join_summary.s_elements = s.size;
        join_summary.rs_elements = m;
    }

    JoinConfig join_config;
    JoinSummary join_summary;
    vj_ctx *ctx = nullptr;
    std::string create_error;
};

#endif  // VJ_JOIN_PROVIDER_HPP
