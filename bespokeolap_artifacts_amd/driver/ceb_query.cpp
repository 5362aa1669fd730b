// ceb_query.cpp — the reference's CEB load() / build() / query() sequence for
// the factorised COUNT queries 1a and 2a (libbolap's ceb_count kernels) and
// the grouped MIN / MAX query 10a (bolap_ceb_q10a).
//
//   bolap_ceb <dir/>  <  requests
//
// load   reads <dir><table>.parquet for the tables of the IMDB schema whose
//        files exist (bespoke_ceb/loader_impl.cpp:9-35; the path is a
//        directory prefix).
// build  keeps every int32 column with one is_null byte per row, its
//        has_nulls flag and the min / max of its non-null values, and every
//        string column as ids into one string pool, -1 for nulls
//        (bespoke_ceb/builder_impl.hpp:97-110, builder_impl.cpp:34-253).  The
//        pool interns tables in the builder's table order and columns in file
//        order, first appearance first; the reference's own ids depend on its
//        thread timing (builder_impl.cpp:407-431), and no COUNT depends on
//        them.  The tables whose order a served query can observe are put
//        in the builder's orders (:432-452): movie_info by (movie_id,
//        info_type_id), cast_info by (movie_id, role_id), name and title by
//        id, with ties in load order (the reference's std::sort leaves them
//        unspecified); 10a emits its rows in the order its loops first reach
//        them.  Columns are copied to the device when a query first reads
//        them.
// query  reads one request per line until an empty line (query_impl.cpp:
//        242-256), runs them in order, prints "N | Execution ms: T" (the
//        timer around run_qN only) and writes resultN.csv: count_star() for
//        1a and 2a (:52-58), the name,info,min,max rows of 10a with the
//        reference's quoting (:68-88, 140-160); an id of the workload
//        without a GPU pipeline here
//        gets the reference's stub file "result / 0" (:45-51) and a note on
//        stderr; an unknown id is an error, as there (:222-224).
#include <arrow/api.h>
#include <arrow/io/file.h>
#include <parquet/arrow/reader.h>

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "bolap.h"

namespace {

void check(int status) {
    if (status != BOLAP_OK) throw std::runtime_error(bolap_last_error());
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool exists(const std::string& path) { return std::ifstream(path).good(); }

std::shared_ptr<arrow::Table> read_parquet(const std::string& path) {
    auto file = arrow::io::ReadableFile::Open(path);
    if (!file.ok()) throw std::runtime_error(path + ": " + file.status().ToString());
    auto reader = parquet::arrow::OpenFile(*file, arrow::default_memory_pool());
    if (!reader.ok()) throw std::runtime_error(path + ": " + reader.status().ToString());
    auto table = (*reader)->ReadTable();
    if (!table.ok()) throw std::runtime_error(path + ": " + table.status().ToString());
    return *table;
}

// ---------------------------------------------------------------------------
// build: string pool, columns, device mirror
// ---------------------------------------------------------------------------
struct Pool {
    std::unordered_map<std::string, int32_t> index;
    std::vector<std::string> values;
    int32_t intern(const std::string& s) {
        auto it = index.find(s);
        if (it != index.end()) return it->second;
        const int32_t id = int32_t(values.size());
        index.emplace(s, id);
        values.push_back(s);
        return id;
    }
    bool find(const std::string& s, int32_t& id) const {
        auto it = index.find(s);
        if (it == index.end()) return false;
        id = it->second;
        return true;
    }
};

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 4);
        hip_check(hipMalloc(&p, bytes), "hipMalloc");
        if (!v.empty())
            hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice),
                      "hipMemcpy");
        return static_cast<T*>(p);
    }
};

struct Column {
    bool str = false;
    std::vector<int32_t> v;     // int32 values, or string ids (-1 for null)
    std::vector<uint8_t> nul;   // 1 = null
    bool has_nulls = false;
    bool any = false;           // a non-null value exists (int32 columns)
    int32_t mn = 0, mx = -1;
    mutable std::shared_ptr<DevBuf> dv, dn;
    const int32_t* dev() const {
        if (!dv) {
            dv = std::make_shared<DevBuf>();
            dv->upload(v);
        }
        return static_cast<const int32_t*>(dv->p);
    }
    const uint8_t* dev_null() const {
        if (!has_nulls) return nullptr;
        if (!dn) {
            dn = std::make_shared<DevBuf>();
            dn->upload(nul);
        }
        return static_cast<const uint8_t*>(dn->p);
    }
    bool null(size_t i) const { return nul[i] != 0; }
};

struct Table {
    int64_t rows = 0;
    std::unordered_map<std::string, Column> cols;
    const Column& at(const std::string& table, const std::string& c) const {
        auto it = cols.find(c);
        if (it == cols.end()) throw std::runtime_error("missing column " + table + "." + c);
        return it->second;
    }
};

struct Db {
    Pool pool;
    std::unordered_map<std::string, Table> tables;
    const Table& at(const std::string& t) const {
        auto it = tables.find(t);
        if (it == tables.end()) throw std::runtime_error("missing table " + t);
        return it->second;
    }
    const Column& col(const std::string& t, const std::string& c) const { return at(t).at(t, c); }
};

void append_strings(const arrow::Array& a, Column& c, Pool& pool) {
    const int64_t n = a.length();
    auto put = [&](int64_t i, const std::string& s) {
        if (a.IsNull(i)) {
            c.v.push_back(-1);
            c.nul.push_back(1);
            c.has_nulls = true;
        } else {
            c.v.push_back(pool.intern(s));
            c.nul.push_back(0);
        }
    };
    if (a.type_id() == arrow::Type::STRING) {
        const auto& s = static_cast<const arrow::StringArray&>(a);
        for (int64_t i = 0; i < n; ++i) put(i, a.IsNull(i) ? std::string() : s.GetString(i));
    } else if (a.type_id() == arrow::Type::LARGE_STRING) {
        const auto& s = static_cast<const arrow::LargeStringArray&>(a);
        for (int64_t i = 0; i < n; ++i) put(i, a.IsNull(i) ? std::string() : s.GetString(i));
    } else {
        const auto& d = static_cast<const arrow::DictionaryArray&>(a);
        Column dict;
        append_strings(*d.dictionary(), dict, pool);
        for (int64_t i = 0; i < n; ++i) {
            if (a.IsNull(i)) {
                c.v.push_back(-1);
                c.nul.push_back(1);
                c.has_nulls = true;
            } else {
                const int32_t id = dict.v[size_t(d.GetValueIndex(i))];
                c.v.push_back(id);
                c.nul.push_back(id < 0 ? 1 : 0);
                c.has_nulls = c.has_nulls || id < 0;
            }
        }
    }
}

// One Parquet table: int32 and string columns (the types the pipelines here
// read); other column types are left out.
Table build_table(const arrow::Table& t, Pool& pool) {
    Table out;
    out.rows = t.num_rows();
    for (int k = 0; k < t.num_columns(); ++k) {
        const auto& name = t.field(k)->name();
        const auto& type = t.field(k)->type();
        const bool is_str = type->id() == arrow::Type::STRING ||
                            type->id() == arrow::Type::LARGE_STRING ||
                            (type->id() == arrow::Type::DICTIONARY &&
                             (static_cast<const arrow::DictionaryType&>(*type).value_type()->id() ==
                                  arrow::Type::STRING ||
                              static_cast<const arrow::DictionaryType&>(*type).value_type()->id() ==
                                  arrow::Type::LARGE_STRING));
        if (!is_str && type->id() != arrow::Type::INT32) continue;
        Column c;
        c.str = is_str;
        c.v.reserve(size_t(out.rows));
        c.nul.reserve(size_t(out.rows));
        for (const auto& chunk : t.column(k)->chunks()) {
            if (is_str) {
                append_strings(*chunk, c, pool);
                continue;
            }
            const auto& a = static_cast<const arrow::Int32Array&>(*chunk);
            for (int64_t i = 0; i < a.length(); ++i) {
                if (a.IsNull(i)) {
                    c.v.push_back(0);
                    c.nul.push_back(1);
                    c.has_nulls = true;
                    continue;
                }
                const int32_t x = a.Value(i);
                c.v.push_back(x);
                c.nul.push_back(0);
                c.mn = c.any ? std::min(c.mn, x) : x;
                c.mx = c.any ? std::max(c.mx, x) : x;
                c.any = true;
            }
        }
        out.cols.emplace(name, std::move(c));
    }
    return out;
}

// ---------------------------------------------------------------------------
// request arguments (bespoke_ceb/args_parser.hpp): IN lists ('a', 'b', ...)
// with '' for a quote inside a quoted value and bare values trimmed, then
// std::quoted scalars
// ---------------------------------------------------------------------------
std::vector<std::string> in_list(std::istringstream& in) {
    char c = 0;
    in >> std::ws >> c;
    if (c != '(') {
        std::ostringstream m;
        m << "Expected '(' at start of IN list, but got '" << c
          << "' (int=" << int(static_cast<unsigned char>(c)) << ")";
        throw std::runtime_error(m.str());
    }
    std::vector<std::string> out;
    for (bool first = true; in >> std::ws; first = false) {
        if (in.peek() == ')') {
            in.get();
            break;
        }
        if (!first) {
            in >> std::ws >> c;
            if (c != ',') throw std::runtime_error("Expected ',' between IN list elements");
        }
        std::string v;
        in >> std::ws;
        if (in.peek() == '\'') {
            in.get();
            for (int ch; (ch = in.get()) != EOF;) {
                if (ch == '\'') {
                    if (in.peek() != '\'') break;
                    in.get();
                }
                v.push_back(char(ch));
            }
        } else {
            while (in && in.peek() != ',' && in.peek() != ')' && in.peek() != EOF)
                v.push_back(char(in.get()));
            const auto b = v.find_first_not_of(" \t\r\n");
            v = b == std::string::npos ? std::string()
                                       : v.substr(b, v.find_last_not_of(" \t\r\n") - b + 1);
        }
        out.push_back(v);
    }
    return out;
}

std::string quoted_arg(std::istringstream& in, const char* who, const char* what) {
    std::string v;
    if (!(in >> std::quoted(v)))
        throw std::runtime_error(std::string(who) + ": failed to parse " + what);
    return v;
}

bool is_null_token(const std::string& v) { return v == "<<NULL>>" || v == "NULL" || v == "null"; }

// Integers of an IN list; null tokens and values std::stoi rejects are
// skipped (query_q1a.cpp:14-28, query_q2a.cpp:32-46).
std::vector<int32_t> int_values(const std::vector<std::string>& vals) {
    std::vector<int32_t> out;
    for (const auto& v : vals) {
        if (is_null_token(v)) continue;
        try {
            out.push_back(int32_t(std::stoi(v)));
        } catch (const std::exception&) {
        }
    }
    return out;
}

// Pool ids of the IN-list strings present in the pool (query_q1a.cpp:30-44).
std::vector<int32_t> string_ids(const Pool& pool, const std::vector<std::string>& vals) {
    std::vector<int32_t> out;
    for (const auto& v : vals) {
        int32_t id;
        if (!is_null_token(v) && pool.find(v, id)) out.push_back(id);
    }
    return out;
}

std::vector<uint8_t> mask_of(const std::vector<int32_t>& ids, size_t size = 0) {
    int32_t mx = -1;
    for (int32_t v : ids) mx = std::max(mx, v);
    std::vector<uint8_t> m(std::max(size, size_t(mx + 1)), 0);
    for (int32_t v : ids)
        if (v >= 0) m[size_t(v)] = 1;
    return m;
}

// id -> 1 for the dimension rows whose string column holds one of `ids`
// (both non-null), over at least `min_size` entries.
std::vector<uint8_t> dim_mask(const Db& db, const char* table, const char* id_col,
                              const char* str_col, const std::set<int32_t>& ids,
                              size_t min_size) {
    const auto& t = db.at(table);
    const auto& id = t.at(table, id_col);
    const auto& s = t.at(table, str_col);
    std::vector<uint8_t> m(min_size, 0);
    for (int64_t i = 0; i < t.rows; ++i) {
        if (id.null(size_t(i)) || s.null(size_t(i)) || id.v[size_t(i)] < 0) continue;
        if (!ids.count(s.v[size_t(i)])) continue;
        const size_t k = size_t(id.v[size_t(i)]);
        if (k >= m.size()) m.resize(k + 1, 0);
        m[k] = 1;
    }
    return m;
}

// id -> 1 for every non-null, non-negative id of a table.
std::vector<uint8_t> valid_ids(const Db& db, const char* table, size_t min_size) {
    const auto& t = db.at(table);
    const auto& id = t.at(table, "id");
    std::vector<uint8_t> m(min_size, 0);
    for (int64_t i = 0; i < t.rows; ++i) {
        if (id.null(size_t(i)) || id.v[size_t(i)] < 0) continue;
        const size_t k = size_t(id.v[size_t(i)]);
        if (k >= m.size()) m.resize(k + 1, 0);
        m[k] = 1;
    }
    return m;
}

// ---------------------------------------------------------------------------
// the count pipelines
// ---------------------------------------------------------------------------
struct Masks {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    bolap_ceb_pred pred(const Column& c, const std::vector<uint8_t>& m) {
        bufs.push_back(std::make_unique<DevBuf>());
        const uint8_t* d = bufs.back()->upload(m);
        return bolap_ceb_pred{c.dev(), c.dev_null(), d, int64_t(m.size())};
    }
};

bolap_ceb_term term_of(const Column& movie, int64_t rows, uint32_t weight) {
    bolap_ceb_term t{};
    t.rows = rows;
    t.movie = movie.dev();
    t.movie_null = movie.dev_null();
    t.weight = weight;
    return t;
}

bool any_set(const std::vector<uint8_t>& m) {
    return std::find(m.begin(), m.end(), uint8_t(1)) != m.end();
}

int64_t count_on_gpu(bolap_ctx* ctx, bolap_ceb_count_args& a) {
    bolap_ceb_count_result r{};
    check(bolap_ceb_count(ctx, &a, &r));
    return r.count;
}

// run_q1a (bespoke_ceb/query_q1a.cpp:156-621): masks as there; COUNT = the sum
// over passing title rows of cast * info1 * info2.
struct Q1a {
    std::vector<std::string> id1, id2, info1, info2, kind, role, gender;
    std::string year1, year2;
};

Q1a parse_1a(const std::string& line) {
    std::istringstream in(line);
    std::string qid;
    if (!(in >> qid)) throw std::runtime_error("Q1a: failed to parse query id");
    Q1a q;
    q.id1 = in_list(in);
    q.id2 = in_list(in);
    q.info1 = in_list(in);
    q.info2 = in_list(in);
    q.kind = in_list(in);
    q.role = in_list(in);
    q.gender = in_list(in);
    q.year1 = quoted_arg(in, "Q1a", "YEAR1");
    q.year2 = quoted_arg(in, "Q1a", "YEAR2");
    return q;
}

std::set<int32_t> as_set(const std::vector<int32_t>& v) { return {v.begin(), v.end()}; }

int64_t run_1a(bolap_ctx* ctx, const Db& db, const Q1a& q) {
    const int year1 = std::stoi(q.year1), year2 = std::stoi(q.year2);
    const auto kind_ids = as_set(string_ids(db.pool, q.kind));
    const auto role_ids = as_set(string_ids(db.pool, q.role));
    const auto gender_ids = as_set(string_ids(db.pool, q.gender));
    auto uniq = [](std::vector<int32_t> v) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return v;
    };
    const auto info1 = uniq(string_ids(db.pool, q.info1));
    const auto info2 = uniq(string_ids(db.pool, q.info2));
    const auto& kt = db.at("kind_type");
    const auto& rt = db.at("role_type");
    const auto& nm = db.at("name");
    const auto& it = db.at("info_type");
    const auto kind_mask = dim_mask(db, "kind_type", "id", "kind", kind_ids, size_t(kt.rows) + 1);
    const auto role_mask = dim_mask(db, "role_type", "id", "role", role_ids, size_t(rt.rows) + 1);
    const auto person_mask = dim_mask(db, "name", "id", "gender", gender_ids, size_t(nm.rows) + 1);
    const size_t type_size = valid_ids(db, "info_type", size_t(it.rows) + 1).size();
    // ID lists over the info_type id range, without an existence check
    // (build_int_mask, query_q1a.cpp:139-152)
    std::vector<uint8_t> type1(type_size, 0), type2(type_size, 0);
    for (int32_t v : int_values(q.id1))
        if (v >= 0 && size_t(v) < type_size) type1[size_t(v)] = 1;
    for (int32_t v : int_values(q.id2))
        if (v >= 0 && size_t(v) < type_size) type2[size_t(v)] = 1;
    const auto value1 = mask_of(info1), value2 = mask_of(info2);
    if (!any_set(kind_mask) || !any_set(role_mask) || !any_set(person_mask) ||
        !any_set(type1) || !any_set(type2) || value1.empty() || value2.empty())
        return 0;  // some factor is zero for every movie
    const auto& title = db.at("title");
    const auto& t_id = title.at("title", "id");
    const auto& t_year = title.at("title", "production_year");
    const auto& mi = db.at("movie_info");
    const auto& ci = db.at("cast_info");
    Masks mk;
    bolap_ceb_count_args a{};
    a.title_rows = title.rows;
    a.t_id = t_id.dev();
    a.t_id_null = t_id.dev_null();
    a.t_year = t_year.dev();
    a.t_year_null = t_year.dev_null();
    a.year_lo = year2;
    a.year_hi = year1;
    a.movie_cap = t_id.any && t_id.mx >= 0 ? t_id.mx + 1 : 0;
    a.n_title_preds = 1;
    a.title_preds[0] = mk.pred(title.at("title", "kind_id"), kind_mask);
    a.n_terms = 3;
    for (int k = 0; k < 2; ++k) {
        auto& t = a.terms[k] = term_of(mi.at("movie_info", "movie_id"), mi.rows, 1);
        t.n_preds = 2;
        t.preds[0] = mk.pred(mi.at("movie_info", "info_type_id"), k ? type2 : type1);
        t.preds[1] = mk.pred(mi.at("movie_info", "info"), k ? value2 : value1);
    }
    auto& c = a.terms[2] = term_of(ci.at("cast_info", "movie_id"), ci.rows, 1);
    c.n_preds = 2;
    c.preds[0] = mk.pred(ci.at("cast_info", "role_id"), role_mask);
    c.preds[1] = mk.pred(ci.at("cast_info", "person_id"), person_mask);
    return count_on_gpu(ctx, a);
}

// run_q2a (bespoke_ceb/query_q2a.cpp:205-503).
struct Q2a : Q1a {};

// IN-list strings resolved as query_q2a.cpp:75-121 and query_q10a.cpp:
// 229-265 do: a value split at its commas is joined back with ", " or ","
// when the joined text is known (`known`); the longest such join starting at
// a value wins.
template <class Known>
std::vector<int32_t> merged_string_filter(const std::vector<std::string>& vals, Known known) {
    std::vector<int32_t> out;
    for (size_t i = 0; i < vals.size();) {
        if (is_null_token(vals[i])) {
            ++i;
            continue;
        }
        std::string with_space = vals[i], without = vals[i];
        int32_t found = -1;
        size_t last = i;
        for (size_t j = i; j < vals.size(); ++j) {
            if (j != i) {
                if (is_null_token(vals[j])) break;
                with_space += ", " + vals[j];
                without += "," + vals[j];
            }
            int32_t id;
            if (known(with_space, id)) found = id, last = j;
            if (known(without, id)) found = id, last = j;
        }
        if (found < 0) {
            ++i;
            continue;
        }
        if (std::find(out.begin(), out.end(), found) == out.end()) out.push_back(found);
        i = last + 1;
    }
    return out;
}

// 2a: a joined text counts when it is a value of the column.
std::vector<int32_t> column_string_filter(const Db& db, const Column& col,
                                          const std::vector<std::string>& vals) {
    std::unordered_set<int32_t> present;
    for (size_t i = 0; i < col.v.size(); ++i)
        if (!col.null(i)) present.insert(col.v[i]);
    return merged_string_filter(vals, [&](const std::string& s, int32_t& id) {
        return db.pool.find(s, id) && present.count(id) != 0;
    });
}

int64_t run_2a(bolap_ctx* ctx, const Db& db, const Q1a& q) {
    const int year1 = std::stoi(q.year1), year2 = std::stoi(q.year2);
    const auto ids1 = int_values(q.id1), ids2 = int_values(q.id2);
    const auto kind_mask = dim_mask(db, "kind_type", "id", "kind",
                                    as_set(string_ids(db.pool, q.kind)), 0);
    const auto role_mask = dim_mask(db, "role_type", "id", "role",
                                    as_set(string_ids(db.pool, q.role)), 0);
    const auto person_mask = dim_mask(db, "name", "id", "gender",
                                      as_set(string_ids(db.pool, q.gender)), 0);
    const auto valid_type = valid_ids(db, "info_type", 0);
    const auto valid_keyword = valid_ids(db, "keyword", 0);
    if (kind_mask.empty() || role_mask.empty() || person_mask.empty() || ids1.empty() ||
        ids2.empty())
        return 0;  // query_q2a.cpp:226-233
    const auto& mi = db.at("movie_info");
    const auto& mi_info = mi.at("movie_info", "info");
    const auto info1 = column_string_filter(db, mi_info, q.info1);
    const auto info2 = column_string_filter(db, mi_info, q.info2);
    if (info1.empty() || info2.empty()) return 0;  // :283-285
    // an info type passes when it is a valid id and in the ID list
    auto type_mask = [&](const std::vector<int32_t>& ids) {
        std::vector<uint8_t> m(valid_type.size(), 0);
        for (int32_t v : ids)
            if (v >= 0 && size_t(v) < m.size() && valid_type[size_t(v)]) m[size_t(v)] = 1;
        return m;
    };
    const auto type1 = type_mask(ids1), type2 = type_mask(ids2);
    const auto& title = db.at("title");
    const auto& t_id = title.at("title", "id");
    const auto& t_year = title.at("title", "production_year");
    const auto& ci = db.at("cast_info");
    const auto& mkw = db.at("movie_keyword");
    if (!any_set(type1) || !any_set(type2) || valid_keyword.empty()) return 0;
    Masks mk;
    bolap_ceb_count_args a{};
    a.title_rows = title.rows;
    a.t_id = t_id.dev();
    a.t_id_null = t_id.dev_null();
    a.t_year = t_year.dev();
    a.t_year_null = t_year.dev_null();
    a.year_lo = year2;
    a.year_hi = year1;
    a.movie_cap = t_id.any && t_id.mx >= 0 ? t_id.mx + 1 : 0;
    a.n_title_preds = 1;
    a.title_preds[0] = mk.pred(title.at("title", "kind_id"), kind_mask);
    a.n_terms = 4;
    for (int k = 0; k < 2; ++k) {
        auto& t = a.terms[k] = term_of(mi.at("movie_info", "movie_id"), mi.rows, 1);
        t.n_preds = 2;
        t.preds[0] = mk.pred(mi.at("movie_info", "info_type_id"), k ? type2 : type1);
        t.preds[1] = mk.pred(mi_info, mask_of(k ? info2 : info1));
    }
    auto& c = a.terms[2] = term_of(ci.at("cast_info", "movie_id"), ci.rows, 1);
    c.n_preds = 2;
    c.preds[0] = mk.pred(ci.at("cast_info", "role_id"), role_mask);
    c.preds[1] = mk.pred(ci.at("cast_info", "person_id"), person_mask);
    auto& w = a.terms[3] = term_of(mkw.at("movie_keyword", "movie_id"), mkw.rows, 1);
    w.n_preds = 1;
    w.preds[0] = mk.pred(mkw.at("movie_keyword", "keyword_id"), valid_keyword);
    return count_on_gpu(ctx, a);
}

// ---------------------------------------------------------------------------
// 10a: SELECT n.name, mi1.info, MIN(t.production_year), MAX(...)
// (bespoke_ceb/query_q10a.cpp:334-1120)
// ---------------------------------------------------------------------------
struct Q10a {
    std::vector<std::string> id, info, kind, role;
    std::string name;
};

Q10a parse_10a(const std::string& line) {  // args_parser.hpp:623-649
    std::istringstream in(line);
    std::string qid;
    if (!(in >> qid)) throw std::runtime_error("Q10a: failed to parse query id");
    Q10a q;
    q.id = in_list(in);
    q.info = in_list(in);
    q.name = quoted_arg(in, "Q10a", "NAME");
    q.kind = in_list(in);
    q.role = in_list(in);
    return q;
}

// id -> 1 for the rows of a dimension table whose string column holds one of
// `ids` (both non-null), over rows + 1 entries; ids outside are dropped
// (build_allowed_id_mask, query_q10a.cpp:183-209).
std::vector<uint8_t> allowed_ids(const Db& db, const char* table, const char* str_col,
                                 const std::vector<int32_t>& ids, bool& any) {
    const auto& t = db.at(table);
    const auto& id = t.at(table, "id");
    const auto& s = t.at(table, str_col);
    const std::unordered_set<int32_t> want(ids.begin(), ids.end());
    std::vector<uint8_t> m(size_t(t.rows) + 1, 0);
    any = false;
    for (int64_t i = 0; i < t.rows; ++i) {
        if (id.null(size_t(i)) || s.null(size_t(i))) continue;
        if (!want.count(s.v[size_t(i)])) continue;
        const int32_t k = id.v[size_t(i)];
        if (k < 0 || size_t(k) >= m.size()) continue;
        m[size_t(k)] = 1;
        any = true;
    }
    return m;
}

// The pool's strings on the device, for ILIKE (uploaded on first use).
struct PoolDev {
    std::unique_ptr<DevBuf> chars, offsets;
    int64_t n = -1;
    void ensure(const Pool& pool) {
        if (n == int64_t(pool.values.size())) return;
        std::vector<char> c;
        std::vector<int64_t> off{0};
        for (const auto& v : pool.values) {
            c.insert(c.end(), v.begin(), v.end());
            off.push_back(int64_t(c.size()));
        }
        chars = std::make_unique<DevBuf>();
        offsets = std::make_unique<DevBuf>();
        chars->upload(c);
        offsets->upload(off);
        n = int64_t(pool.values.size());
    }
};

bolap_ceb_col col_of(const Column& c) { return bolap_ceb_col{c.dev(), c.dev_null()}; }

std::vector<bolap_minmax_row> run_10a(bolap_ctx* ctx, const Db& db, PoolDev& pdev,
                                      const Q10a& q) {
    const std::vector<bolap_minmax_row> none;
    const auto types = int_values(q.id);
    if (types.empty()) return none;  // :352-358
    auto in_pool = [&](const std::string& s, int32_t& id) { return db.pool.find(s, id); };
    const auto kind_f = merged_string_filter(q.kind, in_pool);
    const auto role_f = merged_string_filter(q.role, in_pool);
    const auto info_f = merged_string_filter(q.info, in_pool);
    // a dense index per INFO string (:373-400; any numbering gives the same rows)
    const int64_t pool_n = int64_t(db.pool.values.size());
    std::vector<int32_t> info_dense(size_t(pool_n), -1), info_ids;
    for (int32_t id : info_f) {
        if (id < 0 || id >= pool_n || info_dense[size_t(id)] != -1) continue;
        info_dense[size_t(id)] = int32_t(info_ids.size());
        info_ids.push_back(id);
    }
    bool kind_any = false, role_any = false;
    const auto kind_ok = allowed_ids(db, "kind_type", "kind", kind_f, kind_any);
    if (!kind_any || role_f.empty() || info_f.empty()) return none;  // :402-409
    const auto role_ok = allowed_ids(db, "role_type", "role", role_f, role_any);
    if (!role_any) return none;  // :410-417
    const auto& info_type = db.at("info_type");
    std::vector<uint8_t> type_ok(size_t(info_type.rows) + 1, 0);
    bool type_any = false;
    for (int32_t v : types) {
        if (v < 0 || size_t(v) >= type_ok.size()) continue;
        type_ok[size_t(v)] = 1;
        type_any = true;
    }
    if (!type_any) return none;  // :418-433
    // the NAME pattern (:137-173): quotes stripped, lower-cased; '%' alone
    // matches every name
    std::string pat = q.name;
    if (pat.size() >= 2 && pat.front() == '\'' && pat.back() == '\'')
        pat = pat.substr(1, pat.size() - 2);
    std::string low = pat;
    for (char& c : low) c = char(std::tolower(static_cast<unsigned char>(c)));
    const auto& title = db.at("title");
    const auto& mi = db.at("movie_info");
    const auto& ci = db.at("cast_info");
    const auto& nm = db.at("name");
    DevBuf d_kind, d_role, d_type, d_dense, d_ids, d_name;
    bolap_ceb_q10a_args a{};
    a.title_rows = title.rows;
    a.t_id = col_of(title.at("title", "id"));
    a.t_kind_id = col_of(title.at("title", "kind_id"));
    a.t_year = col_of(title.at("title", "production_year"));
    a.mi_rows = mi.rows;
    a.mi_movie = col_of(mi.at("movie_info", "movie_id"));
    a.mi_type = col_of(mi.at("movie_info", "info_type_id"));
    a.mi_info = col_of(mi.at("movie_info", "info"));
    a.ci_rows = ci.rows;
    a.ci_movie = col_of(ci.at("cast_info", "movie_id"));
    a.ci_person = col_of(ci.at("cast_info", "person_id"));
    a.ci_role = col_of(ci.at("cast_info", "role_id"));
    a.name_rows = nm.rows;
    a.n_id = col_of(nm.at("name", "id"));
    a.n_name = col_of(nm.at("name", "name"));
    a.kind_ok = d_kind.upload(kind_ok);
    a.kind_n = int64_t(kind_ok.size());
    a.role_ok = d_role.upload(role_ok);
    a.role_n = int64_t(role_ok.size());
    a.type_ok = d_type.upload(type_ok);
    a.type_n = int64_t(type_ok.size());
    a.pool_n = pool_n;
    a.info_dense = d_dense.upload(info_dense);
    a.info_ids = d_ids.upload(info_ids);
    a.n_infos = int32_t(info_ids.size());
    if (low != "%") {
        pdev.ensure(db.pool);
        uint8_t* m = d_name.upload(std::vector<uint8_t>(size_t(std::max<int64_t>(pool_n, 1))));
        check(bolap_ilike_mask_device(ctx, static_cast<const char*>(pdev.chars->p),
                                      static_cast<const int64_t*>(pdev.offsets->p), pool_n,
                                      pat.c_str(), m));
        a.name_ok = m;
    }
    std::vector<bolap_minmax_row> rows(4096);
    int64_t n = 0;
    int st = bolap_ceb_q10a(ctx, &a, rows.data(), int64_t(rows.size()), &n, nullptr);
    if (st == BOLAP_ERR_CAPACITY) {
        rows.resize(size_t(n));
        st = bolap_ceb_q10a(ctx, &a, rows.data(), n, &n, nullptr);
    }
    check(st);
    rows.resize(size_t(n));
    return rows;
}

// write_csv_field (query_impl.cpp:68-88): quoted when the text holds a
// comma, quote, newline, carriage return or backslash; quote and backslash
// then take a backslash.
void csv_field(std::ostream& out, const std::string& v) {
    if (v.find_first_of(",\"\n\r\\") == std::string::npos) {
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

void write_10a(const std::string& path, const Db& db, const std::vector<bolap_minmax_row>& rows) {
    std::ofstream out(path);  // query_impl.cpp:140-160
    out << "name,info,min(t.production_year),max(t.production_year)\n";
    for (const auto& r : rows) {
        csv_field(out, db.pool.values[size_t(r.group)]);
        out << ",";
        csv_field(out, db.pool.values[size_t(r.key)]);
        out << ",";
        if (r.has) out << r.min;
        out << ",";
        if (r.has) out << r.max;
        out << "\n";
    }
}

// The builder's row order of one table by int32 columns (builder_impl.cpp:
// 265-340), ties in load order; a missing key column leaves it as loaded.
void sort_table(bolap_ctx* ctx, Table& t, const char* primary, const char* secondary) {
    auto p = t.cols.find(primary);
    if (p == t.cols.end() || p->second.str || t.rows < 2) return;
    const Column* sec = nullptr;
    if (secondary) {
        auto it = t.cols.find(secondary);
        if (it == t.cols.end() || it->second.str) return;
        sec = &it->second;
    }
    DevBuf dp, ds, dord;
    const int32_t* pp = dp.upload(p->second.v);
    const int32_t* sp = sec ? ds.upload(sec->v) : nullptr;
    uint32_t* ord = dord.upload(std::vector<uint32_t>(size_t(t.rows)));
    check(bolap_sort_order_i32(ctx, pp, sp, t.rows, ord));
    std::vector<uint32_t> order(size_t(t.rows));
    hip_check(hipMemcpy(order.data(), ord, order.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    for (auto& [name, c] : t.cols) {
        std::vector<int32_t> v(c.v.size());
        std::vector<uint8_t> nu(c.nul.size());
        for (size_t i = 0; i < order.size(); ++i) {
            v[i] = c.v[order[i]];
            nu[i] = c.nul[order[i]];
        }
        c.v.swap(v);
        c.nul.swap(nu);
    }
}

const char* const kTables[] = {
    "aka_name", "aka_title", "cast_info", "char_name", "comp_cast_type", "company_name",
    "company_type", "complete_cast", "info_type", "keyword", "kind_type", "link_type",
    "movie_companies", "movie_info", "movie_info_idx", "movie_keyword", "movie_link", "name",
    "person_info", "role_type", "title"};

const std::set<std::string> kWorkload = {"1a", "2a", "2b", "2c", "3a", "3b", "4a", "5a",
                                         "6a", "7a", "8a", "9a", "9b", "10a", "11a", "11b"};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: bolap_ceb <dir/>  < requests\n";
        return 2;
    }
    try {
        bolap_ctx* ctx = nullptr;
        check(bolap_ctx_create(0, nullptr, &ctx));
        const std::string dir = argv[1];
        Db db;
        for (const char* name : kTables) {
            const std::string path = dir + name + ".parquet";
            if (exists(path)) db.tables.emplace(name, build_table(*read_parquet(path), db.pool));
        }
        // builder_impl.cpp:432-452, for the tables a served query reads in order
        const struct {
            const char *table, *primary, *secondary;
        } orders[] = {{"movie_info", "movie_id", "info_type_id"},
                      {"cast_info", "movie_id", "role_id"},
                      {"name", "id", nullptr},
                      {"title", "id", nullptr}};
        for (const auto& o : orders) {
            auto it = db.tables.find(o.table);
            if (it != db.tables.end()) sort_table(ctx, it->second, o.primary, o.secondary);
        }
        PoolDev pdev;
        std::vector<std::pair<std::string, std::string>> requests;
        std::string line;
        while (std::getline(std::cin, line)) {
            if (line.empty()) break;
            std::istringstream iss(line);
            std::string id;
            if (!(iss >> id)) continue;
            requests.emplace_back(id, line);
        }
        for (size_t i = 0; i < requests.size(); ++i) {
            const auto& [id, text] = requests[i];
            const std::string csv = "result" + std::to_string(i + 1) + ".csv";
            if (id == "1a" || id == "2a") {
                const Q1a q = parse_1a(text);
                const auto start = std::chrono::steady_clock::now();
                const int64_t n = id == "1a" ? run_1a(ctx, db, q) : run_2a(ctx, db, q);
                const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                                    std::chrono::steady_clock::now() - start)
                                    .count();
                std::cout << (i + 1) << " | Execution ms: " << ms << "\n";
                std::ofstream(csv) << "count_star()\n" << n << "\n";
                continue;
            }
            if (id == "10a") {
                const Q10a q = parse_10a(text);
                const auto start = std::chrono::steady_clock::now();
                const auto rows = run_10a(ctx, db, pdev, q);
                const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                                    std::chrono::steady_clock::now() - start)
                                    .count();
                std::cout << (i + 1) << " | Execution ms: " << ms << "\n";
                write_10a(csv, db, rows);
                continue;
            }
            if (!kWorkload.count(id)) throw std::runtime_error("Unsupported query id: " + id);
            std::cerr << "bolap_ceb: query " << id << " is not served; wrote the stub result\n";
            std::cout << (i + 1) << " | Execution ms: 0\n";
            std::ofstream(csv) << "result\n0\n";
        }
        std::cout.flush();
        bolap_ctx_destroy(ctx);
    } catch (const std::exception& e) {
        std::cerr << "bolap_ceb: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
