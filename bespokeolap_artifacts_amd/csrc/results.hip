// results.hip — host-side result rows and CSV output of the Q1, Q6 and Q18
// pipelines, in the reference's formats.  No device code.
//
// Q1: run_q1's finalisation (bespoke_tpch/query_q1.cpp:234-275) turns the
// exact integer accumulators into doubles (sums / scale, averages / count),
// drops empty groups and orders rows by (returnflag, linestatus) strings;
// write_q1_csv (:283-298) prints fixed notation with two decimals (four for
// avg_disc).  Q6: write_q6_csv (query_q6.cpp:399-408).  Q14: write_q14_csv
// (query_q14.cpp:367-375).  Q17: write_q17_csv (query_q17.cpp:311-322).  Q18: write_q18_csv
// (query_q18.cpp:465-484) with its date format (:41-52) and field quoting
// (:65-79).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

struct Scales {
    double qty, price, disc_price, charge, disc;
};

// Fixed-point scales of the two layouts.  Reference layout: quantity x 100,
// prices in cents, discounted_price in cents, charge in cents x hundredths
// (query_q1.cpp:15-16, 247-255).  Declared-width layout: quantity in units,
// disc_price = price * (100 - disc) in cents x hundredths, charge in cents x
// hundredths x hundredths.
Scales scales_of(int32_t layout) {
    if (layout == BOLAP_Q1_LAYOUT_REF) return Scales{100.0, 100.0, 100.0, 100.0 * 100.0, 100.0};
    return Scales{1.0, 100.0, 100.0 * 100.0, 100.0 * 100.0 * 100.0, 100.0};
}

void civil_from_days(int z, int& y, unsigned& m, unsigned& d) {
    z += 719468;
    const int era = (z >= 0 ? z : z - 146096) / 146097;
    const unsigned doe = unsigned(z - era * 146097);
    const unsigned yoe = (doe - doe / 1460u + doe / 36524u - doe / 146096u) / 365u;
    const unsigned doy = doe - (365u * yoe + yoe / 4u - yoe / 100u);
    const unsigned mp = (5u * doy + 2u) / 153u;
    d = doy - (153u * mp + 2u) / 5u + 1u;
    m = mp < 10u ? mp + 3u : mp - 9u;
    y = int(yoe) + era * 400 + (m <= 2);
}

constexpr int kBaseDays = 8035;  // 1992-01-01

std::string date_string(int32_t offset) {
    int y = 0;
    unsigned m = 0, d = 0;
    civil_from_days(kBaseDays + offset, y, m, d);
    std::ostringstream o;
    o << std::setw(4) << std::setfill('0') << y << '-' << std::setw(2) << std::setfill('0') << m
      << '-' << std::setw(2) << std::setfill('0') << d;
    return o.str();
}

// Always quoted; a double quote inside becomes backslash + quote.
std::string quoted(const char* s) {
    std::string out = "\"";
    for (const char* p = s ? s : ""; *p; ++p) {
        if (*p == '"') out += "\\\"";
        else out += *p;
    }
    out += '"';
    return out;
}

}  // namespace

using namespace bolap;

extern "C" {

int bolap_q1_finalize(const bolap_q1_group* groups, int32_t layout, int32_t n_rf, int32_t n_ls,
                      const char* const* rf_dict, const char* const* ls_dict,
                      bolap_q1_row* rows_out, int32_t* n_out) {
    BOLAP_ARG_CHECK(groups && rf_dict && ls_dict && rows_out && n_out,
                    "bolap_q1_finalize: null argument");
    BOLAP_ARG_CHECK(layout == BOLAP_Q1_LAYOUT_WIDE || layout == BOLAP_Q1_LAYOUT_REF,
                    "bolap_q1_finalize: layout %d", layout);
    BOLAP_ARG_CHECK(n_rf >= 1 && n_ls >= 1 && n_rf * n_ls <= BOLAP_Q1_MAX_GROUPS,
                    "bolap_q1_finalize: %d x %d groups", n_rf, n_ls);
    const Scales sc = scales_of(layout);
    struct Row {
        std::string rf, ls;
        bolap_q1_row r;
    };
    std::vector<Row> rows;
    for (int32_t idx = 0; idx < n_rf * n_ls; ++idx) {
        const bolap_q1_group& g = groups[idx];
        if (g.count == 0) continue;
        Row x;
        x.rf = rf_dict[idx / n_ls] ? rf_dict[idx / n_ls] : "";
        x.ls = ls_dict[idx % n_ls] ? ls_dict[idx % n_ls] : "";
        bolap_q1_row& r = x.r;
        std::memset(&r, 0, sizeof(r));
        std::strncpy(r.returnflag, x.rf.c_str(), sizeof(r.returnflag) - 1);
        std::strncpy(r.linestatus, x.ls.c_str(), sizeof(r.linestatus) - 1);
        const double count = double(g.count);
        const __int128 charge =
            (__int128(g.sum_charge_hi) << 64) | __int128(uint64_t(g.sum_charge_lo));
        r.sum_qty = double(g.sum_qty) / sc.qty;
        r.sum_base_price = double(g.sum_base_price) / sc.price;
        r.sum_disc_price = double(g.sum_disc_price) / sc.disc_price;
        r.sum_charge = double(charge) / sc.charge;
        r.avg_qty = r.sum_qty / count;
        r.avg_price = r.sum_base_price / count;
        r.avg_disc = double(g.sum_disc) / sc.disc / count;
        r.count_order = g.count;
        rows.push_back(std::move(x));
    }
    std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) {
        if (a.rf == b.rf) return a.ls < b.ls;
        return a.rf < b.rf;
    });
    for (size_t i = 0; i < rows.size(); ++i) rows_out[i] = rows[i].r;
    *n_out = int32_t(rows.size());
    return BOLAP_OK;
}

int bolap_q1_write_csv(const char* path, const bolap_q1_row* rows, int32_t n) {
    BOLAP_ARG_CHECK(path && (rows || n == 0) && n >= 0, "bolap_q1_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q1_write_csv: cannot open %s", path);
    out << "l_returnflag,l_linestatus,sum_qty,sum_base_price,sum_disc_price,sum_charge,avg_qty,"
           "avg_price,avg_disc,count_order\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2);
    for (int32_t i = 0; i < n; ++i) {
        const bolap_q1_row& r = rows[i];
        out << r.returnflag << ',' << r.linestatus << ',' << r.sum_qty << ',' << r.sum_base_price
            << ',' << r.sum_disc_price << ',' << r.sum_charge << ',' << r.avg_qty << ','
            << r.avg_price << ',';
        out << std::setprecision(4) << r.avg_disc << std::setprecision(2) << ','
            << r.count_order << "\n";
    }
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

int bolap_q6_write_csv(const char* path, const bolap_q6_result* result) {
    BOLAP_ARG_CHECK(path && result, "bolap_q6_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q6_write_csv: cannot open %s", path);
    out << "revenue\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2) << double(result->revenue) / (100.0 * 100.0) << "\n";
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

int bolap_q14_write_csv(const char* path, const bolap_q14_result* result) {
    BOLAP_ARG_CHECK(path && result, "bolap_q14_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q14_write_csv: cannot open %s", path);
    out << "promo_revenue\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2) << result->promo_revenue << "\n";
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

int bolap_q17_write_csv(const char* path, const bolap_q17_result* result) {
    BOLAP_ARG_CHECK(path && result, "bolap_q17_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q17_write_csv: cannot open %s", path);
    out << "avg_yearly\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2) << double(result->sum_extendedprice) / (7.0 * 100.0) << "\n";
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

int bolap_q3_write_csv(const char* path, const bolap_group_row* rows, int64_t n,
                       const int32_t* shippriority) {
    BOLAP_ARG_CHECK(path && (rows || n == 0) && n >= 0, "bolap_q3_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q3_write_csv: cannot open %s", path);
    out << "l_orderkey,revenue,o_orderdate,o_shippriority\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2);
    for (int64_t i = 0; i < n; ++i) {
        const bolap_group_row& r = rows[i];
        out << r.key << ',' << double(r.revenue) / 100.0 << ',' << date_string(r.day) << ','
            << (shippriority ? shippriority[i] : 0) << "\n";
    }
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

int bolap_q18_write_csv(const char* path, const bolap_q18_row* rows, int64_t n,
                        const char* const* c_name, int64_t n_names) {
    BOLAP_ARG_CHECK(path && (rows || n == 0) && n >= 0 && (c_name || n_names == 0),
                    "bolap_q18_write_csv: bad arguments");
    std::ofstream out(path);
    BOLAP_ARG_CHECK(out.good(), "bolap_q18_write_csv: cannot open %s", path);
    out << "c_name,c_custkey,o_orderkey,o_orderdate,o_totalprice,sum(l_quantity)\n";
    out.setf(std::ios::fixed);
    out << std::setprecision(2);
    for (int64_t i = 0; i < n; ++i) {
        const bolap_q18_row& r = rows[i];
        BOLAP_ARG_CHECK(r.cust_row >= 0 && r.cust_row < n_names,
                        "bolap_q18_write_csv: customer row %d has no name", r.cust_row);
        out << quoted(c_name[r.cust_row]) << ',' << r.custkey << ',' << r.orderkey << ','
            << date_string(r.orderdate) << ',' << double(r.totalprice) / 100.0 << ','
            << double(r.sum_qty) / 100.0 << "\n";
    }
    return out.good() ? BOLAP_OK : BOLAP_ERR_ARG;
}

}  // extern "C"
