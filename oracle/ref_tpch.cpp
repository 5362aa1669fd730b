// Test infrastructure: a C shim around the reference's own TPC-H pipelines.
//
// This translation unit #includes the reference's query_q1.cpp .. query_q22.cpp
// from where they lie (the reference links them into one unit the same way,
// bespoke_tpch/query_impl.cpp:13-56) and compiles them unchanged.  The shim
// adds four things of its own:
//   * a Database filled column by column from caller arrays, field names as
//     in bespoke_tpch/builder_impl.hpp:11-171 ("lineitem.orderkey", ...);
//   * the lineitem shards, formed from those columns as the reference's
//     build() forms them (Q14 reads lineitem only through them);
//   * ref_tpch_run(), which parses one request line, runs run_qN and writes
//     the result with the reference's write_qN_csv, as the reference query()
//     loop does for a served id (query_impl.cpp:161-389);
//   * string columns get all-ones alpha/bigram masks: the masks are only a
//     conservative prefilter in front of the full match (query_q13.cpp:199-211),
//     so all-ones selects the reference's own exact-match path everywhere.
//
// Built only where the reference sources exist, into oracle/_ref/ (see
// bespokeolap_artifacts_amd/build.py:build_ref).  Only tests load it, as a
// checker; the product never links it.  Built a second time with
// REF_TPCH_DROPIN, it is the reference's query unit with the GPU drop-in
// bodies in place of six pipelines, which tests compare against the first.
#ifdef REF_TPCH_DROPIN
// The drop-in variant: run_qN / write_qN_csv of Q1, Q3, Q6, Q14, Q17 and Q18
// come from bespokeolap_artifacts_amd/dropin/tpch_run_q.cpp (the GPU bodies)
// instead of the reference's files; everything else is unchanged.
#include "tpch_run_q.cpp"
#else
#include "query_q1.cpp"
#include "query_q3.cpp"
#include "query_q6.cpp"
#include "query_q14.cpp"
#include "query_q17.cpp"
#include "query_q18.cpp"
#endif
#include "query_q2.cpp"
#include "query_q4.cpp"
#include "query_q5.cpp"
#include "query_q7.cpp"
#include "query_q8.cpp"
#include "query_q9.cpp"
#include "query_q10.cpp"
#include "query_q11.cpp"
#include "query_q12.cpp"
#include "query_q13.cpp"
#include "query_q15.cpp"
#include "query_q16.cpp"
#include "query_q19.cpp"
#include "query_q20.cpp"
#include "query_q21.cpp"
#include "query_q22.cpp"

#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

namespace refshim {

struct Slot {
    size_t width;
    std::function<void(const void*, size_t)> assign;
};

template <class T>
Slot vec(std::vector<T>& v) {
    return {sizeof(T), [&v](const void* p, size_t n) {
                const T* t = static_cast<const T*>(p);
                v.assign(t, t + n);
            }};
}

std::map<std::string, Slot> numeric_slots(Database& d) {
    auto& l = d.lineitem;
    auto& o = d.orders;
    auto& c = d.customer;
    auto& p = d.part;
    auto& s = d.supplier;
    auto& ps = d.partsupp;
    return {
        {"lineitem.orderkey", vec(l.orderkey)},
        {"lineitem.partkey", vec(l.partkey)},
        {"lineitem.suppkey", vec(l.suppkey)},
        {"lineitem.supp_nationkey", vec(l.supp_nationkey)},
        {"lineitem.linenumber", vec(l.linenumber)},
        {"lineitem.shipdate", vec(l.shipdate)},
        {"lineitem.commit_receipt", vec(l.commit_receipt)},
        {"lineitem.extendedprice", vec(l.extendedprice)},
        {"lineitem.discount", vec(l.discount)},
        {"lineitem.tax", vec(l.tax)},
        {"lineitem.quantity", vec(l.quantity)},
        {"lineitem.discounted_price", vec(l.discounted_price)},
        {"lineitem.returnflag_linestatus", vec(l.returnflag_linestatus)},
        {"orders.orderkey", vec(o.orderkey)},
        {"orders.custkey", vec(o.custkey)},
        {"orders.cust_nationkey", vec(o.cust_nationkey)},
        {"orders.orderdate", vec(o.orderdate)},
        {"orders.shippriority", vec(o.shippriority)},
        {"orders.totalprice", vec(o.totalprice)},
        {"orders.orderkey_to_row", vec(o.orderkey_to_row)},
        {"orders.orders_by_customer_offsets", vec(o.orders_by_customer_offsets)},
        {"orders.orders_by_customer_rows", vec(o.orders_by_customer_rows)},
        {"customer.custkey", vec(c.custkey)},
        {"customer.nationkey", vec(c.nationkey)},
        {"customer.nationkey_by_custkey", vec(c.nationkey_by_custkey)},
        {"customer.acctbal", vec(c.acctbal)},
        {"customer.phone_prefix_code", vec(c.phone_prefix_code)},
        {"part.partkey", vec(p.partkey)},
        {"part.size", vec(p.size)},
        {"part.retailprice", vec(p.retailprice)},
        {"supplier.suppkey", vec(s.suppkey)},
        {"supplier.nationkey", vec(s.nationkey)},
        {"supplier.nationkey_by_suppkey", vec(s.nationkey_by_suppkey)},
        {"supplier.acctbal", vec(s.acctbal)},
        {"partsupp.partkey", vec(ps.partkey)},
        {"partsupp.suppkey", vec(ps.suppkey)},
        {"partsupp.availqty", vec(ps.availqty)},
        {"partsupp.supplycost", vec(ps.supplycost)},
        {"partsupp.supplycost_availqty", vec(ps.supplycost_availqty)},
    };
}

DictionaryColumn* dict_slot(Database& d, const std::string& name) {
    static const std::map<std::string, DictionaryColumn* (*)(Database&)> m = {
        {"lineitem.returnflag", [](Database& x) { return &x.lineitem.returnflag; }},
        {"lineitem.linestatus", [](Database& x) { return &x.lineitem.linestatus; }},
        {"lineitem.shipinstruct", [](Database& x) { return &x.lineitem.shipinstruct; }},
        {"lineitem.shipmode", [](Database& x) { return &x.lineitem.shipmode; }},
        {"orders.orderstatus", [](Database& x) { return &x.orders.orderstatus; }},
        {"orders.orderpriority", [](Database& x) { return &x.orders.orderpriority; }},
        {"customer.mktsegment", [](Database& x) { return &x.customer.mktsegment; }},
        {"part.mfgr", [](Database& x) { return &x.part.mfgr; }},
        {"part.brand", [](Database& x) { return &x.part.brand; }},
        {"part.container", [](Database& x) { return &x.part.container; }},
        {"part.type", [](Database& x) { return &x.part.type; }},
    };
    auto it = m.find(name);
    return it == m.end() ? nullptr : it->second(d);
}

StringColumn* string_slot(Database& d, const std::string& name) {
    static const std::map<std::string, StringColumn* (*)(Database&)> m = {
        {"lineitem.comment", [](Database& x) { return &x.lineitem.comment; }},
        {"orders.clerk", [](Database& x) { return &x.orders.clerk; }},
        {"orders.comment", [](Database& x) { return &x.orders.comment; }},
        {"customer.name", [](Database& x) { return &x.customer.name; }},
        {"customer.address", [](Database& x) { return &x.customer.address; }},
        {"customer.phone", [](Database& x) { return &x.customer.phone; }},
        {"customer.comment", [](Database& x) { return &x.customer.comment; }},
        {"part.name", [](Database& x) { return &x.part.name; }},
        {"part.comment", [](Database& x) { return &x.part.comment; }},
        {"supplier.name", [](Database& x) { return &x.supplier.name; }},
        {"supplier.address", [](Database& x) { return &x.supplier.address; }},
        {"supplier.phone", [](Database& x) { return &x.supplier.phone; }},
        {"supplier.comment", [](Database& x) { return &x.supplier.comment; }},
        {"partsupp.comment", [](Database& x) { return &x.partsupp.comment; }},
    };
    auto it = m.find(name);
    return it == m.end() ? nullptr : it->second(d);
}

// (year, month) of a day count since 1970-01-01: Hinnant's civil_from_days.
std::pair<int32_t, int32_t> year_month(int32_t z) {
    z += 719468;
    const int32_t era = (z >= 0 ? z : z - 146096) / 146097;
    const uint32_t doe = static_cast<uint32_t>(z - era * 146097);
    const uint32_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
    const uint32_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
    const uint32_t mp = (5 * doy + 2) / 153;
    const uint32_t m = mp < 10 ? mp + 3 : mp - 9;
    return {static_cast<int32_t>(yoe) + era * 400 + (m <= 2), static_cast<int32_t>(m)};
}

// Lineitem shards as the reference's build() forms them (builder_impl.cpp:
// 855-911): one shard per (ship year-month, discount, quantity / 100, supplier
// nation) in first-appearance order, with min/max zones and the row list.
void build_shards(Database& d) {
    auto& l = d.lineitem;
    l.shards.clear();
    std::unordered_map<int64_t, size_t> index;
    for (uint32_t i = 0; i < l.row_count; ++i) {
        const auto [year, month] = year_month(d.base_date_days + l.shipdate[i]);
        const int16_t nation = l.supp_nationkey[i];
        const uint8_t supp = static_cast<uint8_t>(nation >= 0 ? nation : 255);
        const int64_t key = (static_cast<int64_t>(year * 12 + month - 1) << 24) |
                            (static_cast<int64_t>(l.discount[i]) << 16) |
                            (static_cast<int64_t>(l.quantity[i] / 100) << 8) | supp;
        auto it = index.find(key);
        if (it == index.end()) {
            LineitemShard sh;
            sh.year = year;
            sh.month = month;
            sh.supp_nationkey = supp;
            sh.min_shipdate = sh.max_shipdate = l.shipdate[i];
            sh.min_discount = sh.max_discount = l.discount[i];
            sh.min_quantity = sh.max_quantity = l.quantity[i];
            l.shards.push_back(std::move(sh));
            it = index.emplace(key, l.shards.size() - 1).first;
        }
        auto& sh = l.shards[it->second];
        sh.row_indices.push_back(i);
        sh.min_shipdate = std::min(sh.min_shipdate, l.shipdate[i]);
        sh.max_shipdate = std::max(sh.max_shipdate, l.shipdate[i]);
        sh.min_discount = std::min<int32_t>(sh.min_discount, l.discount[i]);
        sh.max_discount = std::max<int32_t>(sh.max_discount, l.discount[i]);
        sh.min_quantity = std::min(sh.min_quantity, l.quantity[i]);
        sh.max_quantity = std::max(sh.max_quantity, l.quantity[i]);
    }
    for (auto& sh : l.shards) {
        sh.start = sh.row_indices.front();
        sh.end = sh.row_indices.back() + 1;
        sh.contiguous = static_cast<size_t>(sh.end - sh.start) == sh.row_indices.size();
    }
}

void set_error(char* err, size_t cap, const std::string& msg) {
    if (err && cap) {
        const size_t n = std::min(cap - 1, msg.size());
        std::memcpy(err, msg.data(), n);
        err[n] = '\0';
    }
}

int run_line(Database& db, const std::string& line, const std::string& path) {
    std::istringstream iss(line);
    std::string id = "0";
    iss >> id;
    const QueryRequest q{id, line};
#define REF_Q(N)                                                                 \
    if (id == #N) {                                                              \
        auto rows = run_q##N(db, parse_q##N(q));                                 \
        write_q##N##_csv(path, rows);                                            \
        return 0;                                                                \
    }
#define REF_QDB(N)                                                               \
    if (id == #N) {                                                              \
        auto rows = run_q##N(db, parse_q##N(q));                                 \
        write_q##N##_csv(db, path, rows);                                        \
        return 0;                                                                \
    }
    REF_Q(1) REF_Q(2) REF_QDB(3) REF_Q(4) REF_Q(5) REF_Q(6) REF_Q(7) REF_Q(8) REF_Q(9)
    REF_Q(10) REF_Q(11) REF_Q(12) REF_Q(13) REF_Q(14) REF_Q(15) REF_Q(16) REF_Q(17)
    REF_QDB(18) REF_Q(19) REF_Q(20) REF_Q(21) REF_Q(22)
#undef REF_Q
#undef REF_QDB
    return 2;
}

}  // namespace refshim

extern "C" {

void* ref_tpch_db_new(int32_t base_date_days) {
    auto* d = new Database();
    d->base_date_days = base_date_days;
    return d;
}

void ref_tpch_db_free(void* db) {
#ifdef REF_TPCH_DROPIN
    bolap_dropin::release_mirror(*static_cast<Database*>(db));
#endif
    delete static_cast<Database*>(db);
}

// Numeric column `name` from n elements of `width` bytes (checked against the
// field's own element size).  Returns 0, or 1 for an unknown name / width.
int ref_tpch_set(void* db, const char* name, const void* data, uint64_t n, uint64_t width) {
    auto& d = *static_cast<Database*>(db);
    auto slots = refshim::numeric_slots(d);
    auto it = slots.find(name);
    if (it == slots.end() || it->second.width != width) return 1;
    it->second.assign(data, static_cast<size_t>(n));
    return 0;
}

// orders.lineitem_ranges from n (start, end) pairs.
int ref_tpch_set_ranges(void* db, const uint32_t* start, const uint32_t* end, uint64_t n) {
    auto& r = static_cast<Database*>(db)->orders.lineitem_ranges;
    r.resize(static_cast<size_t>(n));
    for (size_t i = 0; i < r.size(); ++i) r[i] = LineitemTable::OrderRange{start[i], end[i]};
    return 0;
}

// Dictionary column: n codes and n_values NUL-terminated strings back to back.
int ref_tpch_set_dict(void* db, const char* name, const uint16_t* codes, uint64_t n,
                      const char* values, uint64_t n_values) {
    auto* col = refshim::dict_slot(*static_cast<Database*>(db), name);
    if (!col) return 1;
    col->codes.assign(codes, codes + n);
    col->dictionary.clear();
    for (uint64_t i = 0; i < n_values; ++i) {
        col->dictionary.emplace_back(values);
        values += col->dictionary.back().size() + 1;
    }
    return 0;
}

// String column: n rows, offsets[n + 1] into `data`.
int ref_tpch_set_strings(void* db, const char* name, const char* data, const uint32_t* offsets,
                         uint64_t n) {
    auto* col = refshim::string_slot(*static_cast<Database*>(db), name);
    if (!col) return 1;
    col->offsets.assign(offsets, offsets + n + 1);
    col->data.assign(data, n ? offsets[n] : 0);
    col->alpha_mask.assign(static_cast<size_t>(n), 0xffffffffu);
    col->bigram_mask.assign(static_cast<size_t>(n), ~0ull);
    return 0;
}

void ref_tpch_add_nation(void* db, int32_t key, int32_t region, const char* name) {
    auto& nt = static_cast<Database*>(db)->nation;
    NationTable::Row row;
    row.nationkey = key;
    row.regionkey = region;
    row.name = name;
    nt.name_to_key.emplace(row.name, key);
    nt.nationkey_to_row.emplace(key, static_cast<uint32_t>(nt.rows.size()));
    nt.rows.push_back(std::move(row));
}

void ref_tpch_add_region(void* db, int32_t key, const char* name) {
    auto& rt = static_cast<Database*>(db)->region;
    RegionTable::Row row;
    row.regionkey = key;
    row.name = name;
    rt.name_to_key.emplace(row.name, key);
    rt.regionkey_to_row.emplace(key, static_cast<uint32_t>(rt.rows.size()));
    rt.rows.push_back(std::move(row));
}

// Row counts, lineitem.orderkey_sorted and the lineitem shards, once every
// column is set.
void ref_tpch_finish(void* db, int orderkey_sorted) {
    auto& d = *static_cast<Database*>(db);
    d.lineitem.row_count = d.lineitem.orderkey.size();
    d.lineitem.orderkey_sorted = orderkey_sorted != 0;
    refshim::build_shards(d);
    d.orders.row_count = d.orders.orderkey.size();
    d.customer.row_count = d.customer.custkey.size();
    d.part.row_count = d.part.partkey.size();
    d.supplier.row_count = d.supplier.suppkey.size();
    d.partsupp.row_count = d.partsupp.partkey.size();
}

// Runs one request line and writes its CSV to csv_path.  0 = done, 2 = an id
// the reference does not serve, 1 = the reference threw (message in err).
int ref_tpch_run(void* db, const char* line, const char* csv_path, char* err, uint64_t err_cap) {
    try {
        return refshim::run_line(*static_cast<Database*>(db), line, csv_path);
    } catch (const std::exception& e) {
        refshim::set_error(err, static_cast<size_t>(err_cap), e.what());
        return 1;
    }
}

}  // extern "C"
