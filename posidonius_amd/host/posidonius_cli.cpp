// posidonius_cli.cpp -- `posidonius-hip start|resume`, the command-line host
// of the reference (src/main.rs:10-190) on top of the C ABI.
//
//   posidonius-hip start  <case.json> <snapshot.json> <history.bin> [-l seconds] [-s]
//   posidonius-hip resume <snapshot.json> <history.bin> [-l seconds] [-s]
//                         [--historic-snapshot-period days] [--recovery-snapshot-period days]
//                         [--time-limit days]
//   posidonius-hip convert <in> <out>       (JSON <-> bincode, by extension)
//   common options: [--device N] [--history-capacity N]
//
// Files:
//   * case / recovery snapshot: the WHFast document the reference's Python
//     package writes and its snapshots hold (src/integrator/output.rs:28-54,
//     178-279), as JSON when the file name ends in ".json" and as bincode
//     1.x otherwise (bincode.hpp); a snapshot is the input document with the
//     integrator and particle state updated.  Before a snapshot is
//     replaced, the old one is kept as <stem>.<YYYYMMDD>T<AM|PM>.bin, one per
//     twelve hours (output.rs:35-41);
//   * history: the reference's binary records, 156 bytes per body and
//     snapshot (output.rs:57-172, bincode little-endian):
//       f64 time, f64 time_step, i32 id, f64 x[3], spin[3], v[3], mass,
//       radius, radius_of_gyration_2, f64 love_number,
//       scaled_dissipation_factor, lag_angle, denergy_dt, migration_timescale.
//
// The device takes the historic snapshots inside the step (pos_history_*);
// the host replays the time accumulation of WHFast::iterate (t += dt) to
// decide how many steps to launch before the next recovery snapshot, the
// completion of the run or a full history buffer.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <limits>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/posidonius_hip.h"
#include "bincode.hpp"
#include "json.hpp"

using posjson::Value;

namespace {

constexpr long long MAX_PARTICLES = 10;   // src/constants.rs:3, the key base of the pair-dependent map

constexpr int HISTORY_RECORD_BYTES = 156;   // output.rs:58-62

std::string now_utc() {
    const std::time_t t = std::time(nullptr);
    char buf[32];
    std::strftime(buf, sizeof buf, "%Y.%m.%d %H:%M:%S", std::gmtime(&t));
    return buf;
}
void info(const std::string &msg) { std::printf("[INFO %s UTC] %s\n", now_utc().c_str(), msg.c_str()); std::fflush(stdout); }
[[noreturn]] void die(const std::string &msg) {
    std::fprintf(stderr, "[PANIC %s UTC] %s\n", now_utc().c_str(), msg.c_str());
    std::exit(1);
}
bool exists(const std::string &p) { struct stat st; return ::stat(p.c_str(), &st) == 0; }

std::string read_file(const std::string &p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) die("cannot open '" + p + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// The extension decides the format, as in the reference (output.rs:48, 186).
bool is_json_path(const std::string &p) {
    const size_t slash = p.find_last_of('/');
    const size_t dot = p.find_last_of('.');
    return dot != std::string::npos && (slash == std::string::npos || dot > slash) && p.substr(dot + 1) == "json";
}

Value read_document(const std::string &p) {
    const std::string bytes = read_file(p);
    return is_json_path(p) ? posjson::parse(bytes) : posbin::decode(bytes);
}

// Path::with_extension(ext): the file name's last extension replaced.
std::string with_extension(const std::string &p, const std::string &ext) {
    const size_t slash = p.find_last_of('/');
    const size_t dot = p.find_last_of('.');
    const bool has_ext = dot != std::string::npos && (slash == std::string::npos || dot > slash + 1);
    return (has_ext ? p.substr(0, dot) : p) + "." + ext;
}

// The backup name of an existing snapshot: one per twelve hours, UTC
// (output.rs:35-41, time's "[year][month][day]T[period]" format).
std::string backup_path(const std::string &p) {
    const std::time_t t = std::time(nullptr);
    const std::tm *g = std::gmtime(&t);
    char buf[32];
    std::snprintf(buf, sizeof buf, "%04d%02d%02dT%s.bin", g->tm_year + 1900, g->tm_mon + 1, g->tm_mday,
                  g->tm_hour < 12 ? "AM" : "PM");
    return with_extension(p, buf);
}

// serde externally tagged enum: "Tag" or {"Tag": payload}
std::string enum_tag(const Value &v, const Value **payload = nullptr) {
    if (payload) *payload = nullptr;
    if (v.is_string()) return v.str();
    if (v.is_object() && v.members.size() == 1) {
        if (payload) *payload = &v.members[0].second;
        return v.members[0].first;
    }
    throw std::runtime_error("unexpected enum value");
}

int evolution_code(const Value &v) {
    const Value *payload;
    const std::string tag = enum_tag(v, &payload);
    if (tag == "NonEvolving") return POS_EVO_NON_EVOLVING;
    if (tag == "Leconte2011") return POS_EVO_LECONTE2011;
    if (tag == "Baraffe1998") return POS_EVO_BARAFFE1998;
    if (tag == "Baraffe2015") return POS_EVO_BARAFFE2015;
    if (tag == "BolmontMathis2016") return POS_EVO_BOLMONT_MATHIS2016;
    if (tag == "GalletBolmont2017") return POS_EVO_GALLET_BOLMONT2017;
    if (tag == "LeconteChabrier2013")
        return (payload && payload->kind == Value::Bool && payload->b) ? POS_EVO_LECONTE_CHABRIER2013_DYN
                                                                       : POS_EVO_LECONTE_CHABRIER2013;
    throw std::runtime_error("unknown evolution type " + tag);
}

// obj[key] = v, adding the member when the document lacks it
void put_member(Value &obj, const std::string &key, Value v) {
    if (Value *m = obj.find(key)) *m = std::move(v);
    else obj.members.emplace_back(key, std::move(v));
}
double num_or(const Value &obj, const char *key, double dflt) {
    const Value *v = obj.find(key);
    return v ? v->num() : dflt;
}

void axes(const Value &v, double *out) { out[0] = v.at("x").num(); out[1] = v.at("y").num(); out[2] = v.at("z").num(); }
void set_axes(Value &v, const double *in) {
    v.at("x") = Value::number(in[0]);
    v.at("y") = Value::number(in[1]);
    v.at("z") = Value::number(in[2]);
}

// One WHFast case as the C ABI's structure of arrays (one system).
struct Case {
    Value doc;
    int n = 0;
    std::vector<int32_t> body_offset, coordinates, gr_impl;
    std::vector<uint32_t> effects, roles;
    std::vector<double> time_step, time_limit, current_time, hist_period, last_hist;
    std::vector<int64_t> current_iteration;
    std::vector<double> mass, radius, rg2, inertia, position, velocity, spin, angular_momentum, helio_velocity;
    std::vector<double> sigma, sigma_scale, k2, k2f, wind_k, wind_sat, err_v, err_L;
    std::vector<int32_t> evo_type, evo_offset, evo_left, ids;
    std::vector<double> tab_time, tab_radius, tab_rg2, tab_love, tab_itq;
    std::vector<double> disk_properties = std::vector<double>(6, 0.);   // the disk host's DiskProperties
    std::vector<double> pair_factor;   // per body: its pair-dependent map entry with the host, NaN = none
    std::vector<double> roche;         // per body: POS_MAX_BODIES Roche radii (universe.roche_radiuses)
    int host = 0;                      // most massive body

    explicit Case(Value d) : doc(std::move(d)) {
        if (!doc.find("alternative_coordinates_type")) throw std::runtime_error("not a WHFast case");
        const Value &u = doc.at("universe");
        n = (int)u.at("n_particles").integer();
        if (n < 2 || n > POS_MAX_BODIES) throw std::runtime_error("unsupported number of particles");
        body_offset = {0, n};
        const std::string ct = doc.at("alternative_coordinates_type").str();
        coordinates = {ct == "Jacobi" ? POS_COORD_JACOBI : ct == "WHDS" ? POS_COORD_WHDS : POS_COORD_DEMOCRATIC_HELIOCENTRIC};
        const Value &ce = u.at("consider_effects");
        uint32_t eff = 0;
        const std::pair<const char *, uint32_t> bits[] = {
            {"tides", POS_EFFECT_TIDES}, {"rotational_flattening", POS_EFFECT_ROTATIONAL_FLATTENING},
            {"general_relativity", POS_EFFECT_GENERAL_RELATIVITY}, {"disk", POS_EFFECT_DISK},
            {"wind", POS_EFFECT_WIND}, {"evolution", POS_EFFECT_EVOLUTION}};
        for (const auto &b : bits)
            if (ce.at(b.first).b) eff |= b.second;
        effects = {eff};
        const std::string gri = enum_tag(u.at("general_relativity_implementation"));
        gr_impl = {gri == "Kidder1995" ? POS_GR_KIDDER1995 : gri == "Anderson1975" ? POS_GR_ANDERSON1975
                   : gri == "Newhall1983" ? POS_GR_NEWHALL1983 : POS_GR_DISABLED};
        time_step = {doc.at("time_step").num()};
        time_limit = {u.at("time_limit").num()};
        current_time = {doc.at("current_time").num()};
        hist_period = {doc.at("historic_snapshot_period").num()};
        last_hist = {doc.at("last_historic_snapshot_time").num()};
        current_iteration = {doc.at("current_iteration").integer()};
        const Value &ps = u.at("particles");
        const Value *evs = u.find("particles_evolvers");
        evo_offset.push_back(0);
        for (int i = 0; i < n; ++i) {
            const Value &p = ps[i];
            ids.push_back(p.find("id") ? (int32_t)p.at("id").integer() : i);
            mass.push_back(p.at("mass").num());
            radius.push_back(p.at("radius").num());
            rg2.push_back(p.at("radius_of_gyration_2").num());
            inertia.push_back(p.at("moment_of_inertia").num());
            double a[3];
            axes(p.at("inertial_position"), a); position.insert(position.end(), a, a + 3);
            axes(p.at("inertial_velocity"), a); velocity.insert(velocity.end(), a, a + 3);
            axes(p.at("spin"), a); spin.insert(spin.end(), a, a + 3);
            axes(p.at("angular_momentum"), a); angular_momentum.insert(angular_momentum.end(), a, a + 3);
            axes(p.at("heliocentric_velocity"), a); helio_velocity.insert(helio_velocity.end(), a, a + 3);
            uint32_t r = 0;
            double s = 0., sc = 0., k = 0., kf = 0., wk = 0., ws = 0.;
            const Value *pl;
            std::string tag = enum_tag(p.at("tides").at("effect"), &pl);
            if (tag == "CentralBody" || tag == "OrbitingBody") {
                const Value *params;
                if (enum_tag(*pl, &params) != "ConstantTimeLag") throw std::runtime_error("only the constant-time-lag tidal model is supported");
                r |= tag == "CentralBody" ? POS_ROLE_TIDES_CENTRAL : POS_ROLE_TIDES_ORBITING;
                s = params->at("dissipation_factor").num();
                sc = params->at("dissipation_factor_scale").num();
                k = params->at("love_number").num();
            }
            tag = enum_tag(p.at("rotational_flattening").at("effect"), &pl);
            if (tag == "CentralBody" || tag == "OrbitingBody") {
                const Value *params;
                if (enum_tag(*pl, &params) != "OblateSpheroid") throw std::runtime_error("only the oblate-spheroid flattening model is supported");
                r |= tag == "CentralBody" ? POS_ROLE_FLAT_CENTRAL : POS_ROLE_FLAT_ORBITING;
                kf = params->at("love_number").num();
            }
            tag = enum_tag(p.at("general_relativity").at("effect"), &pl);
            if (tag == "CentralBody" && !(pl && pl->is_string() && pl->str() == "Disabled")) r |= POS_ROLE_GR_CENTRAL;
            else if (tag == "OrbitingBody") r |= POS_ROLE_GR_ORBITING;
            if (const Value *disk = p.find("disk")) {   // disk.rs:44-49
                const Value *props;
                tag = enum_tag(disk->at("effect"), &props);
                if (tag == "CentralBody") {
                    r |= POS_ROLE_DISK_CENTRAL;
                    const char *keys[6] = {"inner_edge_distance", "outer_edge_distance", "lifetime", "alpha",
                                           "surface_density_normalization", "mean_molecular_weight"};
                    for (int q = 0; q < 6; ++q) disk_properties[q] = props->at(keys[q]).num();
                } else if (tag == "OrbitingBody") {
                    r |= POS_ROLE_DISK_ORBITING;
                }
            }
            const Value &wind = p.at("wind");
            if (wind.at("effect").is_string() && wind.at("effect").str() == "Interaction") {
                r |= POS_ROLE_WIND;
                wk = wind.at("parameters").at("input").at("k_factor").num();
                ws = wind.at("parameters").at("input").at("rotation_saturation").num();
            }
            roles.push_back(r);
            sigma.push_back(s); sigma_scale.push_back(sc); k2.push_back(k); k2f.push_back(kf);
            wind_k.push_back(wk); wind_sat.push_back(ws);
            axes(doc.at("inertial_velocity_errors")[i], a); err_v.insert(err_v.end(), a, a + 3);
            axes(doc.at("particle_angular_momentum_errors")[i], a); err_L.insert(err_L.end(), a, a + 3);
            evo_type.push_back(evolution_code(p.at("evolution")));
            const Value *ev = (evs && evs->is_array() && (size_t)i < evs->items.size()) ? &(*evs)[i] : nullptr;
            if (ev && evolution_code(ev->at("evolution")) != POS_EVO_NON_EVOLVING) {
                const Value &t = ev->at("time");
                const size_t len = t.items.size();
                for (size_t j = 0; j < len; ++j) {
                    tab_time.push_back(t[j].num());
                    tab_radius.push_back(ev->at("radius")[j].num());
                    const Value &rg = ev->at("radius_of_gyration_2");
                    const Value &lv = ev->at("love_number");
                    const Value *iq = ev->find("inverse_tidal_q_factor");
                    tab_rg2.push_back(rg.items.size() == len ? rg[j].num() : 0.);
                    tab_love.push_back(lv.items.size() == len ? lv[j].num() : 0.);
                    tab_itq.push_back(iq && iq->items.size() == len ? (*iq)[j].num() : 0.);
                }
                evo_left.push_back((int32_t)ev->at("left_index").integer());
                evo_offset.push_back(evo_offset.back() + (int32_t)len);
            } else {
                evo_left.push_back(0);
                evo_offset.push_back(evo_offset.back());
            }
        }
        read_pair_map(u);
        // universe.roche_radiuses: row i of an n x n block (universe.rs:177-196)
        roche.assign((size_t)n * POS_MAX_BODIES, 0.);
        if (const Value *rr = u.find("roche_radiuses"))
            if (rr->is_array() && rr->items.size() >= (size_t)n * n)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j) roche[(size_t)i * POS_MAX_BODIES + j] = (*rr)[(size_t)i * n + j].num();
    }

    // the map entries id * MAX_PARTICLES + host_id (constant_time_lag.rs:188-194)
    void read_pair_map(const Value &u) {
        pair_factor.assign(n, std::numeric_limits<double>::quiet_NaN());
        double mmax = 0.;
        for (int i = 0; i < n; ++i)
            if (mass[i] > mmax) { mmax = mass[i]; host = i; }
        const Value *map = u.find("pair_dependent_scaled_dissipation_factor");
        if (!map || !map->is_object()) return;
        for (const auto &m : map->members) {
            const long long key = std::stoll(m.first);
            const long long owner = key / MAX_PARTICLES, dep = key % MAX_PARTICLES;
            if (dep != ids[host]) continue;
            for (int i = 0; i < n; ++i)
                if (i != host && ids[i] == owner) pair_factor[i] = m.second.num();
        }
    }

    pos_ensemble_desc desc() const {
        pos_ensemble_desc d{};
        d.n_systems = 1;
        d.body_offset = body_offset.data();
        d.coordinates = coordinates.data();
        d.effects = effects.data();
        d.gr_implementation = gr_impl.data();
        d.time_step = time_step.data();
        d.time_limit = time_limit.data();
        d.current_time = current_time.data();
        d.historic_snapshot_period = hist_period.data();
        d.last_historic_snapshot_time = last_hist.data();
        d.current_iteration = current_iteration.data();
        d.mass = mass.data();
        d.radius = radius.data();
        d.radius_of_gyration_2 = rg2.data();
        d.moment_of_inertia = inertia.data();
        d.position = position.data();
        d.velocity = velocity.data();
        d.spin = spin.data();
        d.angular_momentum = angular_momentum.data();
        d.heliocentric_velocity = helio_velocity.data();
        d.roles = roles.data();
        d.tides_dissipation_factor = sigma.data();
        d.tides_dissipation_factor_scale = sigma_scale.data();
        d.tides_love_number = k2.data();
        d.flattening_love_number = k2f.data();
        d.wind_k_factor = wind_k.data();
        d.wind_rotation_saturation = wind_sat.data();
        d.velocity_error = err_v.data();
        d.angular_momentum_error = err_L.data();
        d.evolution_type = evo_type.data();
        d.evolver_offset = evo_offset.data();
        d.evolver_time = tab_time.data();
        d.evolver_radius = tab_radius.data();
        d.evolver_radius_of_gyration_2 = tab_rg2.data();
        d.evolver_love_number = tab_love.data();
        d.evolver_left_index = evo_left.data();
        d.evolver_inverse_tidal_q_factor = tab_itq.data();
        d.disk_properties = disk_properties.data();
        d.pair_dissipation_factor = pair_factor.data();
        d.roche_radius = roche.data();
        return d;
    }
};

void check(int rc, const char *what) {
    if (rc != POS_OK) die(std::string(what) + ": " + pos_last_error());
}

// Recovery snapshot: the input document with the current state written in.
void write_recovery_snapshot(const std::string &path, Case &c, pos_ens *ens, double last_recovery,
                             long long n_historic, double hist_period, double recovery_period) {
    const int n = c.n;
    std::vector<double> x(3 * n), v(3 * n), L(3 * n), s(3 * n), ev(3 * n), eL(3 * n), R(n), rg2(n), I(n), t(1), lh(1);
    std::vector<int64_t> it(1);
    std::vector<int32_t> left(n), tw(1);
    std::vector<double> pairs(n), roche((size_t)n * POS_MAX_BODIES);
    pos_state st{};
    st.position = x.data(); st.velocity = v.data(); st.angular_momentum = L.data(); st.spin = s.data();
    st.velocity_error = ev.data(); st.angular_momentum_error = eL.data(); st.radius = R.data();
    st.radius_of_gyration_2 = rg2.data(); st.moment_of_inertia = I.data(); st.current_time = t.data();
    st.last_historic_snapshot_time = lh.data(); st.current_iteration = it.data();
    st.evolver_left_index = left.data(); st.timestep_warning = tw.data(); st.pair_dissipation_factor = pairs.data();
    st.roche_radius = roche.data();
    check(pos_get_state(ens, &st), "pos_get_state");
    Value &d = c.doc;
    put_member(d, "current_time", Value::number(t[0]));
    put_member(d, "current_iteration", Value::integer_value(it[0]));
    put_member(d, "last_historic_snapshot_time", Value::number(lh[0]));
    put_member(d, "last_recovery_snapshot_time", Value::number(last_recovery));
    put_member(d, "n_historic_snapshots", Value::integer_value(n_historic));
    put_member(d, "historic_snapshot_period", Value::number(hist_period));
    put_member(d, "recovery_snapshot_period", Value::number(recovery_period));
    put_member(d, "timestep_warning", Value::integer_value(tw[0]));
    put_member(d, "hash", Value::integer_value(0));   // the reference hashes its Debug text; no counterpart here
    Value &u = d.at("universe");
    u.at("time_limit") = Value::number(c.time_limit[0]);
    Value &ps = u.at("particles");
    Value *evs = u.find("particles_evolvers");
    for (int i = 0; i < n; ++i) {
        Value &p = ps[i];
        // every particle carries its id, its index at creation (universe.rs:119)
        put_member(p, "id", Value::integer_value(c.ids[i]));
        set_axes(p.at("inertial_position"), &x[3 * i]);
        set_axes(p.at("inertial_velocity"), &v[3 * i]);
        set_axes(p.at("angular_momentum"), &L[3 * i]);
        set_axes(p.at("spin"), &s[3 * i]);
        p.at("radius") = Value::number(R[i]);
        p.at("radius_of_gyration_2") = Value::number(rg2[i]);
        p.at("moment_of_inertia") = Value::number(I[i]);
        const double zero[3] = {0., 0., 0.};
        set_axes(p.at("heliocentric_velocity"), zero);   // zero for the host after every step
        set_axes(d.at("inertial_velocity_errors")[i], &ev[3 * i]);
        set_axes(d.at("particle_angular_momentum_errors")[i], &eL[3 * i]);
        if (evs && evs->is_array() && (size_t)i < evs->items.size() && (*evs)[i].find("left_index"))
            (*evs)[i].at("left_index") = Value::integer_value(left[i]);
    }
    // the orbiting bodies' pair-dependent map entries (the host's entries
    // are recomputed before every use and are not written)
    Value map;
    map.kind = Value::Object;
    for (int i = 0; i < n; ++i)
        if (!std::isnan(pairs[i]))
            map.members.emplace_back(std::to_string((long long)c.ids[i] * MAX_PARTICLES + c.ids[c.host]),
                                     Value::number(pairs[i]));
    put_member(u, "pair_dependent_scaled_dissipation_factor", std::move(map));
    {   // the Roche radii computed at t = 0, row i of an n x n block in a
        // MAX_PARTICLES^2 array (universe.rs:62, 177-196)
        Value rr;
        rr.kind = Value::Array;
        const size_t len = std::max<size_t>((size_t)(MAX_PARTICLES * MAX_PARTICLES), (size_t)n * n);
        for (size_t k = 0; k < len; ++k) {
            const size_t i = k / n, j = k % n;
            rr.items.push_back(Value::number(k < (size_t)n * n ? roche[i * POS_MAX_BODIES + j] : 0.));
        }
        put_member(u, "roche_radiuses", std::move(rr));
    }
    std::string out;
    if (is_json_path(path)) {
        posjson::dump(d, out, 2);
    } else {
        try {
            out = posbin::encode(d);
        } catch (const std::exception &e) {
            die(std::string("cannot encode the bincode snapshot: ") + e.what());
        }
    }
    if (exists(path) && std::rename(path.c_str(), backup_path(path).c_str()) != 0) die("cannot back up '" + path + "'");
    std::ofstream f(path, std::ios::binary);
    if (!f) die("cannot write '" + path + "'");
    f << out;
}

template <class T> void put(std::string &buf, T v) { buf.append(reinterpret_cast<const char *>(&v), sizeof(T)); }

void usage() {
    std::fprintf(stderr,
                 "usage: posidonius-hip start <case.json> <snapshot.json> <history.bin> [-l seconds] [-s]\n"
                 "       posidonius-hip resume <snapshot.json> <history.bin> [-l seconds] [-s]\n"
                 "              [--historic-snapshot-period days] [--recovery-snapshot-period days] [--time-limit days]\n"
                 "       common: [--device N] [--history-capacity N]\n"
                 "       posidonius-hip convert <in> <out>   (a case or snapshot between JSON and bincode)\n"
                 "  a file name ending in .json is JSON, any other is the reference's bincode layout\n");
    std::exit(2);
}

// `convert`: one document from one encoding to the other (no device work).
int convert(const std::string &in, const std::string &out) {
    Value doc;
    try {
        doc = read_document(in);
    } catch (const std::exception &e) {
        die(std::string("cannot read '") + in + "': " + e.what());
    }
    std::string bytes;
    try {
        if (is_json_path(out)) posjson::dump(doc, bytes, 2);
        else bytes = posbin::encode(doc);
    } catch (const std::exception &e) {
        die(std::string("cannot encode '") + out + "': " + e.what());
    }
    std::ofstream f(out, std::ios::binary);
    if (!f || !(f << bytes)) die("cannot write '" + out + "'");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const auto wall0 = std::chrono::steady_clock::now();
    if (argc < 2) usage();
    const std::string cmd = argv[1];
    if (cmd == "convert") {
        if (argc != 4) usage();
        return convert(argv[2], argv[3]);
    }
    const bool resume = cmd == "resume";
    if (!resume && cmd != "start") usage();
    std::vector<std::string> pos;
    long long limit_s = 0;
    bool silent = false;
    double new_hist = -1., new_rec = -1., new_tl = -1.;
    int device = 0, capacity = 256;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string { if (i + 1 >= argc) usage(); return argv[++i]; };
        if (a == "-l" || a == "--limit") limit_s = std::stoll(next());
        else if (a == "-s" || a == "--silent") silent = true;
        else if (resume && a == "--historic-snapshot-period") new_hist = std::stod(next());
        else if (resume && a == "--recovery-snapshot-period") new_rec = std::stod(next());
        else if (resume && a == "--time-limit") new_tl = std::stod(next());
        else if (a == "--device") device = std::stoi(next());
        else if (a == "--history-capacity") capacity = std::max(1, std::stoi(next()));
        else if (!a.empty() && a[0] == '-') usage();
        else pos.push_back(a);
    }
    if (pos.size() != (resume ? 2u : 3u)) usage();
    const std::string first = pos[0];
    const std::string snapshot = resume ? pos[0] : pos[1];
    const std::string history = resume ? pos[1] : pos[2];
    if (!exists(first)) die(resume ? "It was not possible to resume the simulation" : "It was not possible to start the simulation");
    if (!resume && exists(snapshot)) die("File '" + snapshot + "' already exists.");
    if (!resume && exists(history)) die("File '" + history + "' already exists.");

    std::unique_ptr<Case> cp;
    try {
        cp.reset(new Case(read_document(first)));
    } catch (const std::exception &e) {
        die(std::string("cannot read '") + first + "': " + e.what());
    }
    Case &c = *cp;
    info(std::string(c.current_time[0] == 0. ? "Created new simulation based on '" : "Restored previous simulation from '") + first + "'.");
    double hist_period = c.hist_period[0];
    double rec_period = num_or(c.doc, "recovery_snapshot_period", 1e300);
    double last_rec = num_or(c.doc, "last_recovery_snapshot_time", -1.);
    long long n_historic = c.doc.find("n_historic_snapshots") ? c.doc.at("n_historic_snapshots").integer() : 0;
    // set_snapshot_periods / set_time_limit (whfast.rs:187-224)
    if (new_hist > 0. && new_hist != hist_period) {
        info("The historic snapshot period changed from " + std::to_string(hist_period) + " to " + std::to_string(new_hist) + " days");
        hist_period = c.hist_period[0] = new_hist;
    }
    if (new_rec > 0. && new_rec != rec_period) {
        info("The recovery snapshot period changed from " + std::to_string(rec_period) + " to " + std::to_string(new_rec) + " days");
        rec_period = new_rec;
    }
    if (new_tl > 0. && new_tl != c.time_limit[0]) {
        if (new_tl < c.time_limit[0] && new_tl < c.current_time[0]) die("Your new time limit is smaller than the current time");
        info("The time limit changed from " + std::to_string(c.time_limit[0]) + " to " + std::to_string(new_tl) + " days");
        c.time_limit[0] = new_tl;
    }

    // history file: keep exactly what the snapshot accounts for (output.rs:64-87)
    const long long expected = n_historic * HISTORY_RECORD_BYTES * (long long)c.n;
    {
        std::FILE *f = std::fopen(history.c_str(), "ab");
        if (!f) die("File error: cannot open '" + history + "'");
        std::fclose(f);
        struct stat st;
        ::stat(history.c_str(), &st);
        if ((long long)st.st_size < expected)
            die("Historic snapshots do not contain all the expected history (" + std::to_string(expected) +
                " bytes) as indicated by the recovery snapshot");
        if (::truncate(history.c_str(), expected) != 0) die("cannot truncate '" + history + "'");
    }
    std::FILE *hist = std::fopen(history.c_str(), "ab");
    if (!hist) die("cannot open '" + history + "'");

    const pos_ensemble_desc desc = c.desc();
    pos_options opts{};
    opts.device = device;
    opts.mapping = POS_MAPPING_PACKED;
    opts.steps_per_launch = 256;
    opts.use_graph = 0;   // long launches: direct launches measure faster than graph replays
    pos_ens *ens = nullptr;
    check(pos_create(&desc, &opts, &ens), "pos_create");
    check(pos_history_enable(ens, capacity), "pos_history_enable");

    // replay of WHFast's bookkeeping on the host
    double t = c.current_time[0], last_hist = c.last_hist[0];
    const double dt = c.time_step[0];
    const double tl = c.time_limit[0];
    std::vector<pos_history_record> recs((size_t)capacity * c.n);
    bool done = false;
    while (!done) {
        // steps of this launch: up to the first step that ends the run, asks for
        // a recovery snapshot, or would overflow the history buffer
        int64_t k = 0;
        int snaps = 0;
        bool want_recovery = false, completes = false;
        // with an execution-time limit the reference writes no periodic
        // recovery snapshot (main.rs:158-176) and checks the clock every
        // step; launches are kept short enough to check it often
        const int64_t max_chunk = limit_s > 0 ? (1 << 14) : (1 << 20);
        while (k < max_chunk) {
            const bool first_trigger = last_hist < 0.;
            const bool hist_trigger = first_trigger || last_hist + hist_period <= t;
            const bool rec_trigger = last_rec + rec_period <= t;
            if (hist_trigger) {
                if (snaps == capacity) break;
                ++snaps;
                last_hist = first_trigger ? 0. : last_hist + hist_period;
            }
            t = t + dt;
            ++k;
            if (t + dt > tl) { completes = true; break; }
            if (limit_s == 0 && (first_trigger || rec_trigger)) { want_recovery = true; break; }
        }
        check(pos_step(ens, k), "pos_step");
        int64_t nrec = 0;
        check(pos_history_drain(ens, recs.data(), (int64_t)recs.size(), &nrec), "pos_history_drain");
        std::string buf;
        for (int64_t r = 0; r < nrec; ++r) {
            const pos_history_record &h = recs[(size_t)r];
            put(buf, h.time); put(buf, h.time_step); put(buf, (int32_t)c.ids[h.body]);
            for (int q = 0; q < 3; ++q) put(buf, h.position[q]);
            for (int q = 0; q < 3; ++q) put(buf, h.spin[q]);
            for (int q = 0; q < 3; ++q) put(buf, h.velocity[q]);
            put(buf, h.mass); put(buf, h.radius); put(buf, h.radius_of_gyration_2);
            put(buf, h.love_number); put(buf, h.scaled_dissipation_factor); put(buf, h.lag_angle);
            put(buf, h.denergy_dt); put(buf, h.migration_timescale);
        }
        if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), hist) != buf.size()) die("history write failed");
        n_historic += nrec / c.n;
        if (!silent && nrec > 0)
            std::printf("Year: %.0f (%.1e) | Time step: %.3f days                    \r", t / 365.25, t / 365.25, dt);
        uint32_t status = 0;
        double t_dev = 0.;
        pos_state st{};
        st.status = &status;
        st.current_time = &t_dev;
        check(pos_get_state(ens, &st), "pos_get_state");
        if (t_dev != t) die("host and device clocks disagree");
        if (status & POS_STATUS_FATAL) {
            std::fflush(hist);
            die(std::string("the simulation stopped on a fatal condition (status ") + std::to_string(status) + ")");
        }
        if (completes) {
            std::printf("\n");
            info("Simulation completed '" + first + "'.");
            done = true;
            break;
        }
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
        if (limit_s > 0 && elapsed >= (double)limit_s) {
            std::fflush(hist);
            last_rec = t;
            write_recovery_snapshot(snapshot, c, ens, last_rec, n_historic, hist_period, rec_period);
            info("Reached execution time limit before simulation completion");
            done = true;
        } else if (want_recovery) {
            std::fflush(hist);
            last_rec = t;
            write_recovery_snapshot(snapshot, c, ens, last_rec, n_historic, hist_period, rec_period);
        }
    }
    std::fclose(hist);
    pos_destroy(ens);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    info(std::string(resume ? "Execution time since last resume: " : "Execution time: ") + std::to_string(wall) + " seconds");
    return 0;
}
