// A caller written the way the reference's join benchmark drives a join
// (src/join_benchmark.cu:209-275): JoinConfig -> JoinProvider -> db_tables ->
// join(r, s, rs) -> get_join_summary() -> rs.free(). It compiles unchanged
// against include/vj/join_provider.hpp and runs on libvj.so.
//
// usage: reference_caller DIR COLUMNS BUCKETS ALGORITHM
//   reads  DIR/r_values.bin DIR/r_pks.bin DIR/s_values.bin DIR/s_pks.bin
//          (uint64, row-major key columns and one primary key per row)
//   writes DIR/rs_values.bin DIR/rs_pks.bin      (device-table join)
//          DIR/rs_host_values.bin DIR/rs_host_pks.bin (host-table join)
//          DIR/summary.txt
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "vj/join_provider.hpp"

static std::vector<column_t> read_u64(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return {};
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<column_t> v((size_t)n / sizeof(column_t));
    f.read(reinterpret_cast<char *>(v.data()), n);
    return v;
}

static void write_u64(const std::string &path, const column_t *p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(column_t)));
}

static db_table upload(const std::vector<column_t> &values, const std::vector<column_t> &pks, int columns) {
    db_table t(columns, pks.size(), true);
    if (!pks.empty()) {
        (void)hipMemcpy(t.column_values, values.data(), values.size() * sizeof(column_t), hipMemcpyHostToDevice);
        (void)hipMemcpy(t.primary_keys, pks.data(), pks.size() * sizeof(column_t), hipMemcpyHostToDevice);
    }
    return t;
}

static void save(const std::string &dir, const std::string &tag, const db_table &rs) {
    db_table h = rs.gpu ? rs.copy(false) : rs;
    write_u64(dir + "/" + tag + "_values.bin", h.column_values, h.size * h.column_count);
    write_u64(dir + "/" + tag + "_pks.bin", h.primary_keys, h.size);
    if (rs.gpu) h.free();
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::cerr << "usage: reference_caller DIR COLUMNS BUCKETS ALGORITHM" << std::endl;
        return 2;
    }
    const std::string dir = argv[1];
    const int columns = std::atoi(argv[2]);

    JoinConfig join_config;
    join_config.buckets = std::atoi(argv[3]);
    join_config.profile_enabled = true;
    join_config.devices = 1;
    join_config.tasks_p_device = 4;
    join_config.hash_config.algorithm = argv[4];

    const std::vector<column_t> rv = read_u64(dir + "/r_values.bin"), rp = read_u64(dir + "/r_pks.bin");
    const std::vector<column_t> sv = read_u64(dir + "/s_values.bin"), sp = read_u64(dir + "/s_pks.bin");
    if (rv.size() != rp.size() * columns || sv.size() != sp.size() * columns) {
        std::cerr << "input sizes do not match the column count" << std::endl;
        return 2;
    }

    JoinSummary total_join_summary;
    {
        JoinProvider join_provider(join_config);
        db_table r_table = upload(rv, rp, columns);
        db_table s_table = upload(sv, sp, columns);
        db_table rs_table;

        auto join_status = join_provider.join(r_table, s_table, rs_table);
        if (join_status.has_failed()) {
            std::cout << "Join failed " << join_status.message << std::endl;
            return 1;
        }
        total_join_summary += join_provider.get_join_summary();
        if (rs_table.column_count != 2 || !rs_table.gpu || !rs_table.data_owner) {
            std::cout << "unexpected rs table shape" << std::endl;
            return 1;
        }
        save(dir, "rs", rs_table);

        // host tables take the reference's CPU branch of join()
        db_table r_copy_table = r_table.copy(false);
        db_table s_copy_table = s_table.copy(false);
        db_table rs_host_table;
        join_status = join_provider.join(r_copy_table, s_copy_table, rs_host_table);
        if (join_status.has_failed() || rs_host_table.gpu) {
            std::cout << "Host join failed " << join_status.message << std::endl;
            return 1;
        }
        save(dir, "rs_host", rs_host_table);

        r_copy_table.free();
        s_copy_table.free();
        rs_host_table.free();
        r_table.free();
        s_table.free();
        rs_table.free();
    }
    std::ofstream f(dir + "/summary.txt");
    f << "rs_elements " << total_join_summary.rs_elements << "\n"
      << "join_tuples_p_second " << total_join_summary.join_tuples_p_second << "\n"
      << "probe_tuples_p_second " << total_join_summary.probe_tuples_p_second << "\n"
      << "status " << total_join_summary.join_status.message << "\n";
    total_join_summary.print();
    return 0;
}
