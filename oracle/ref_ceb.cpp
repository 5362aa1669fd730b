// Test infrastructure: a C shim around the reference's own CEB pipelines.
//
// This translation unit #includes the reference's bespoke_ceb/query_q*.cpp
// from where they lie (its query_impl.cpp includes them the same way,
// bespoke_ceb/query_impl.cpp:11-42) and compiles them unchanged.  The shim
// adds what the reference's build() and query() would otherwise provide:
//   * a Database filled column by column from caller arrays: i32 columns
//     with their is_null bytes, has_nulls and non-null min / max, string
//     columns interned into the StringPool in row order with -1 for nulls
//     (bespoke_ceb/builder_impl.cpp:34-253), then StringPool::finalize();
//   * the builder's table orders (builder_impl.cpp:432-452): movie_info by
//     (movie_id, info_type_id), cast_info by (movie_id, role_id),
//     movie_companies by movie_id, movie_keyword by keyword_id, person_info
//     by info_type_id, name and title by id, with their sorted_by flags.  The
//     reference's std::sort leaves equal keys in any order; this shim keeps
//     input order (a stable sort), one of the orders it may produce;
//   * ref_ceb_run(): one request line through parse_qN and run_qN, written as
//     the reference's query() writes it (query_impl.cpp:52-187): count_star()
//     files for the COUNT queries and the string-pool rows of 3b, 9a, 9b,
//     10a, 11a and 11b with its write_csv_field quoting.
//
// Built only where the reference sources exist, into oracle/_ref/ (see
// bespokeolap_artifacts_amd/build.py:build_ref).  Only tests load it, as a
// checker; the product never links it.
#include "query_q1a.cpp"
#include "query_q2a.cpp"
#include "query_q2b.cpp"
#include "query_q2c.cpp"
#include "query_q3a.cpp"
#include "query_q3b.cpp"
#include "query_q4a.cpp"
#include "query_q5a.cpp"
#include "query_q6a.cpp"
#include "query_q7a.cpp"
#include "query_q8a.cpp"
#include "query_q9a.cpp"
#include "query_q9b.cpp"
#include "query_q10a.cpp"
#include "query_q11a.cpp"
#include "query_q11b.cpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>

namespace refceb {

void set_error(char* err, size_t cap, const std::string& msg) {
    if (err && cap) {
        const size_t n = std::min(cap - 1, msg.size());
        std::memcpy(err, msg.data(), n);
        err[n] = '\0';
    }
}

template <class T>
void reorder(std::vector<T>& v, const std::vector<uint32_t>& order) {
    if (v.size() != order.size()) return;
    std::vector<T> out(order.size());
    for (size_t i = 0; i < order.size(); ++i) out[i] = v[order[i]];
    v.swap(out);
}

void sort_table(TableData& t, const char* primary, const char* secondary) {
    auto p = t.columns.find(primary);
    if (p == t.columns.end() || p->second.type != ColumnType::Int32) return;
    const ColumnData* sec = nullptr;
    if (secondary) {
        auto it = t.columns.find(secondary);
        if (it == t.columns.end() || it->second.type != ColumnType::Int32) return;
        sec = &it->second;
    }
    const auto& pk = p->second.i32;
    std::vector<uint32_t> order(static_cast<size_t>(t.row_count));
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (pk[a] != pk[b]) return pk[a] < pk[b];
        return sec ? sec->i32[a] < sec->i32[b] : false;
    });
    for (auto& e : t.columns) {
        auto& c = e.second;
        reorder(c.i32, order);
        reorder(c.str_ids, order);
        reorder(c.is_null, order);
    }
    t.sorted_by_movie_id = true;
}

void write_field(std::ofstream& out, const std::string& v) {
    bool quote = false;
    for (char c : v)
        if (c == ',' || c == '"' || c == '\n' || c == '\r' || c == '\\') quote = true;
    if (!quote) {
        out << v;
        return;
    }
    out << '"';
    for (char c : v) {
        if (c == '"' || c == '\\') out << '\\';
        out << c;
    }
    out << '"';
}

int run_line(const Database& db, const std::string& line, const std::string& path) {
    std::istringstream iss(line);
    std::string id = "0";
    iss >> id;
    const QueryRequest q{id, line};
    std::ofstream out;
    auto count = [&](int64_t n) {
        out.open(path);
        out << "count_star()\n" << n << "\n";
        return 0;
    };
    if (id == "1a") return count(run_q1a(db, parse_q1a(q)));
    if (id == "2a") return count(run_q2a(db, parse_q2a(q)));
    if (id == "2b") return count(run_q2b(db, parse_q2b(q)));
    if (id == "2c") return count(run_q2c(db, parse_q2c(q)));
    if (id == "3a") return count(run_q3a(db, parse_q3a(q)));
    if (id == "4a") return count(run_q4a(db, parse_q4a(q)));
    if (id == "5a") {
        const auto args = parse_q5a(q);
        prepare_q5a(db);
        return count(run_q5a(db, args));
    }
    if (id == "6a") return count(run_q6a(db, parse_q6a(q)));
    if (id == "7a") return count(run_q7a(db, parse_q7a(q)));
    if (id == "8a") return count(run_q8a(db, parse_q8a(q)));
    const auto& pool = db.string_pool;
    if (id == "10a") {
        const auto rows = run_q10a(db, parse_q10a(q));
        out.open(path);
        out << "name,info,min(t.production_year),max(t.production_year)\n";
        for (const auto& r : rows) {
            write_field(out, pool.get(r.person_name_id));
            out << ",";
            write_field(out, pool.get(r.movie_info_id));
            out << ",";
            if (r.has_year) out << r.min_year;
            out << ",";
            if (r.has_year) out << r.max_year;
            out << "\n";
        }
        return 0;
    }
    if (id == "9a") {
        const auto rows = run_q9a(db, parse_q9a(q));
        out.open(path);
        out << "info,info_1,count_star()\n";
        for (const auto& r : rows) {
            write_field(out, pool.get(r.movie_info_id));
            out << ",";
            write_field(out, pool.get(r.person_info_id));
            out << "," << r.count << "\n";
        }
        return 0;
    }
    if (id == "9b") {
        const auto rows = run_q9b(db, parse_q9b(q));
        out.open(path);
        out << "info,name,count_star()\n";
        for (const auto& r : rows) {
            write_field(out, pool.get(r.movie_info_id));
            out << ",";
            write_field(out, pool.get(r.person_name_id));
            out << "," << r.count << "\n";
        }
        return 0;
    }
    if (id == "3b") {
        const auto rows = run_q3b(db, parse_q3b(q));
        out.open(path);
        out << "title,name,name_1,count_star()\n";
        for (const auto& r : rows) {
            write_field(out, pool.get(r.title_id));
            out << ",";
            write_field(out, pool.get(r.person_name_id));
            out << ",";
            write_field(out, pool.get(r.company_name_id));
            out << "," << r.count << "\n";
        }
        return 0;
    }
    auto grouped = [&](const auto& rows) {
        out.open(path);
        out << "gender,role,name,count_star()\n";
        for (const auto& r : rows) {
            if (r.gender_id >= 0) write_field(out, pool.get(r.gender_id));
            out << ",";
            write_field(out, pool.get(r.role_id));
            out << ",";
            write_field(out, pool.get(r.company_name_id));
            out << "," << r.count << "\n";
        }
        return 0;
    };
    if (id == "11a") return grouped(run_q11a(db, parse_q11a(q)));
    if (id == "11b") return grouped(run_q11b(db, parse_q11b(q)));
    return 2;
}

}  // namespace refceb

extern "C" {

void* ref_ceb_db_new() { return new Database(); }

void ref_ceb_db_free(void* db) { delete static_cast<Database*>(db); }

// An int32 column (is_null may be NULL: no nulls).
void ref_ceb_set_i32(void* db, const char* table, const char* column, const int32_t* values,
                     const uint8_t* is_null, int64_t n) {
    auto& t = static_cast<Database*>(db)->tables[table];
    t.row_count = n;
    ColumnData c;
    c.type = ColumnType::Int32;
    c.i32.assign(values, values + n);
    c.is_null.assign(static_cast<size_t>(n), 0);
    bool any = false;
    for (int64_t i = 0; i < n; ++i) {
        if (is_null && is_null[i]) {
            c.is_null[i] = 1;
            c.i32[i] = 0;
            c.has_nulls = true;
            continue;
        }
        if (!any) {
            c.i32_min = c.i32_max = values[i];
            any = true;
        } else {
            c.i32_min = std::min(c.i32_min, values[i]);
            c.i32_max = std::max(c.i32_max, values[i]);
        }
    }
    t.columns[column] = std::move(c);
}

// A string column: n rows, offsets[n + 1] into chars, interned in row order.
void ref_ceb_set_str(void* db, const char* table, const char* column, const char* chars,
                     const int64_t* offsets, const uint8_t* is_null, int64_t n) {
    auto& d = *static_cast<Database*>(db);
    auto& t = d.tables[table];
    t.row_count = n;
    ColumnData c;
    c.type = ColumnType::String;
    c.str_ids.resize(static_cast<size_t>(n));
    c.is_null.assign(static_cast<size_t>(n), 0);
    for (int64_t i = 0; i < n; ++i) {
        if (is_null && is_null[i]) {
            c.is_null[i] = 1;
            c.str_ids[i] = -1;
            c.has_nulls = true;
            continue;
        }
        c.str_ids[i] = d.string_pool.intern(
            std::string(chars + offsets[i], static_cast<size_t>(offsets[i + 1] - offsets[i])));
    }
    t.columns[column] = std::move(c);
}

// Finalises the string pool and applies the builder's table orders.
void ref_ceb_finish(void* db) {
    auto& d = *static_cast<Database*>(db);
    d.string_pool.finalize();
    auto sort = [&](const char* name, const char* p, const char* s) {
        auto it = d.tables.find(name);
        if (it != d.tables.end()) refceb::sort_table(it->second, p, s);
        return it == d.tables.end() ? nullptr : &it->second;
    };
    if (auto* t = sort("movie_info", "movie_id", "info_type_id")) t->sorted_by_info_type_id = true;
    if (auto* t = sort("cast_info", "movie_id", "role_id")) t->sorted_by_role_id = true;
    sort("movie_companies", "movie_id", nullptr);
    if (auto* t = sort("movie_keyword", "keyword_id", nullptr)) {
        t->sorted_by_movie_id = false;
        t->sorted_by_keyword_id = true;
    }
    if (auto* t = sort("person_info", "info_type_id", nullptr)) {
        t->sorted_by_movie_id = false;
        t->sorted_by_info_type_id = true;
    }
    sort("name", "id", nullptr);
    sort("title", "id", nullptr);
}

// Runs one COUNT request (1a, 2a, 2b, 2c, 3a, 4a, 6a, 7a, 8a) and returns its
// count and the time of run_qN alone (steady_clock, parse outside), as the
// reference's query() times it (query_impl.cpp:262-268).  0 = done, 2 = not a
// COUNT id, 1 = the reference threw (message in err).
int ref_ceb_count_timed(void* db, const char* line, int64_t* count, double* ms, char* err,
                        uint64_t err_cap) {
    try {
        const auto& d = *static_cast<Database*>(db);
        std::istringstream iss(line);
        std::string id = "0";
        iss >> id;
        const QueryRequest q{id, line};
        auto timed = [&](auto&& run) {
            const auto t0 = std::chrono::steady_clock::now();
            *count = run();
            *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                      .count();
            return 0;
        };
#define REF_COUNT(N)                                                                   \
    if (id == #N) {                                                                    \
        const auto args = parse_q##N(q);                                               \
        return timed([&] { return run_q##N(d, args); });                              \
    }
        REF_COUNT(1a) REF_COUNT(2a) REF_COUNT(2b) REF_COUNT(2c) REF_COUNT(3a) REF_COUNT(4a)
        REF_COUNT(6a) REF_COUNT(7a) REF_COUNT(8a)
#undef REF_COUNT
        return 2;
    } catch (const std::exception& e) {
        refceb::set_error(err, static_cast<size_t>(err_cap), e.what());
        return 1;
    }
}

// Runs one request line and writes its CSV to csv_path.  0 = done, 2 = an id
// the reference does not serve, 1 = the reference threw (message in err).
int ref_ceb_run(void* db, const char* line, const char* csv_path, char* err, uint64_t err_cap) {
    try {
        return refceb::run_line(*static_cast<Database*>(db), line, csv_path);
    } catch (const std::exception& e) {
        refceb::set_error(err, static_cast<size_t>(err_cap), e.what());
        return 1;
    }
}

}  // extern "C"
