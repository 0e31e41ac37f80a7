// bincode.hpp -- the reference's binary recovery snapshot, as an encoding of
// the same document the JSON snapshot holds.
//
// The reference writes a snapshot whose name does not end in ".json" with
// bincode 1.x's default options (src/integrator/output.rs:28-54: little
// endian, fixed-width integers): an f64 is its 8 bytes, a usize or u64 8
// bytes, a bool 1 byte, an enum its u32 variant index followed by the
// variant's payload, a Vec and a HashMap a u64 length followed by the
// elements (key, value) and a fixed array ([T; N], serde_big_array) its N
// elements with no length.  A struct is its fields in declaration order; the
// names appear nowhere, so the encoder needs the declaration order of every
// struct of the snapshot, which the schema below lists (WHFast,
// whfast.rs:98-130; Universe and its host structs, universe.rs:17-63;
// Particle, particle.rs:16-52; the effect structs of
// src/effects/{tides,rotational_flattening}/common.rs, general_relativity.rs,
// wind.rs, disk.rs, evolution.rs).  The JSON form of the same structs
// (serde_json, or the reference's Python writer) names every field, so a
// document converts both ways by walking the schema.
//
// Only the constant-time-lag tidal model and the oblate-spheroid flattening
// model are listed (the integration step supports no other); a snapshot that
// uses another model is rejected when read.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"

namespace posbin {

using posjson::Value;

// Type expressions: f64, u64, bool, S:<struct>, E:<enum>, A<n>:<type>,
// V:<type>, M (HashMap<usize, f64>).
struct Schema {
    std::map<std::string, std::vector<std::pair<std::string, std::string>>> structs;
    std::map<std::string, std::vector<std::pair<std::string, std::string>>> enums;   // variant, payload ("" = unit)

    Schema() {
        auto S = [&](const char *name, std::vector<std::pair<std::string, std::string>> fields) { structs[name] = std::move(fields); };
        auto E = [&](const char *name, std::vector<std::pair<std::string, std::string>> variants) { enums[name] = std::move(variants); };
        S("Axes", {{"x", "f64"}, {"y", "f64"}, {"z", "f64"}});
        S("WHFast", {{"time_step", "f64"}, {"half_time_step", "f64"}, {"universe", "S:Universe"}, {"current_time", "f64"},
                     {"current_iteration", "u64"}, {"recovery_snapshot_period", "f64"}, {"historic_snapshot_period", "f64"},
                     {"last_recovery_snapshot_time", "f64"}, {"last_historic_snapshot_time", "f64"},
                     {"n_historic_snapshots", "u64"}, {"hash", "u64"},
                     {"particles_alternative_coordinates", "A10:S:AlternativeCoordinates"},
                     {"alternative_coordinates_type", "E:CoordinatesType"}, {"timestep_warning", "u64"},
                     {"inertial_velocity_errors", "A10:S:Axes"}, {"particle_angular_momentum_errors", "A10:S:Axes"}});
        S("AlternativeCoordinates", {{"mass", "f64"}, {"mass_g", "f64"}, {"position", "S:Axes"}, {"velocity", "S:Axes"},
                                     {"acceleration", "S:Axes"}});
        E("CoordinatesType", {{"Jacobi", ""}, {"DemocraticHeliocentric", ""}, {"WHDS", ""}});
        S("Universe", {{"initial_time", "f64"}, {"time_limit", "f64"}, {"particles", "A10:S:Particle"},
                       {"particles_evolvers", "V:S:Evolver"}, {"n_particles", "u64"}, {"consider_effects", "S:ConsiderEffects"},
                       {"general_relativity_implementation", "E:GeneralRelativityImplementation"}, {"hosts", "S:Hosts"},
                       {"pair_dependent_scaled_dissipation_factor", "M"}, {"roche_radiuses", "A100:f64"}});
        S("HostIndices", {{"most_massive", "u64"}, {"tides", "u64"}, {"rotational_flattening", "u64"},
                          {"general_relativity", "u64"}, {"disk", "u64"}});
        S("HostMostMassive", {{"all", "bool"}, {"general_relativity", "bool"}, {"tides", "bool"},
                              {"rotational_flattening", "bool"}, {"disk", "bool"}});
        S("Hosts", {{"index", "S:HostIndices"}, {"most_massive", "S:HostMostMassive"}});
        S("ConsiderEffects", {{"tides", "bool"}, {"rotational_flattening", "bool"}, {"general_relativity", "bool"},
                              {"disk", "bool"}, {"wind", "bool"}, {"evolution", "bool"}});
        S("Particle", {{"id", "u64"}, {"mass", "f64"}, {"mass_g", "f64"}, {"radius", "f64"},
                       {"inertial_position", "S:Axes"}, {"inertial_velocity", "S:Axes"}, {"inertial_acceleration", "S:Axes"},
                       {"inertial_additional_acceleration", "S:Axes"}, {"heliocentric_position", "S:Axes"},
                       {"heliocentric_velocity", "S:Axes"}, {"heliocentric_distance", "f64"},
                       {"heliocentric_radial_velocity", "f64"}, {"heliocentric_norm_velocity_vector", "f64"},
                       {"heliocentric_norm_velocity_vector_2", "f64"}, {"spin", "S:Axes"}, {"norm_spin_vector_2", "f64"},
                       {"angular_momentum", "S:Axes"}, {"dangular_momentum_dt", "S:Axes"}, {"radius_of_gyration_2", "f64"},
                       {"moment_of_inertia", "f64"}, {"reference", "E:Reference"}, {"tides", "S:Tides"},
                       {"rotational_flattening", "S:RotationalFlattening"}, {"general_relativity", "S:GeneralRelativity"},
                       {"wind", "S:Wind"}, {"disk", "S:Disk"}, {"evolution", "E:EvolutionType"}});
        E("Reference", {{"MostMassiveParticle", ""}, {"Particle", "u64"}});
        // tides (tides/common.rs:41-100, constant_time_lag.rs:12-18)
        S("Tides", {{"effect", "E:TidesEffect"}, {"parameters", "S:TidesParameters"}, {"coordinates", "S:Coordinates"}});
        E("TidesEffect", {{"CentralBody", "E:TidalModel"}, {"OrbitingBody", "E:TidalModel"}, {"Disabled", ""}});
        E("TidalModel", {{"ConstantTimeLag", "S:ConstantTimeLagParameters"}, {"CreepCoplanar", "!"}, {"Kaula", "!"}});
        S("ConstantTimeLagParameters", {{"dissipation_factor", "f64"}, {"dissipation_factor_scale", "f64"}, {"love_number", "f64"}});
        S("TidesParameters", {{"internal", "S:TidesInternal"}, {"output", "S:Output2"}});
        S("TidesInternal", {{"distance", "f64"}, {"radial_velocity", "f64"}, {"scaled_dissipation_factor", "f64"},
                            {"scalar_product_of_vector_position_with_stellar_spin", "f64"},
                            {"scalar_product_of_vector_position_with_planetary_spin", "f64"},
                            {"orthogonal_component_of_the_tidal_force_due_to_stellar_tide", "f64"},
                            {"orthogonal_component_of_the_tidal_force_due_to_planetary_tide", "f64"},
                            {"radial_component_of_the_tidal_force", "f64"},
                            {"radial_component_of_the_tidal_force_dissipative_part_when_star_as_point_mass", "f64"},
                            {"shape", "S:Axes"}, {"denergy_dt", "f64"}, {"lag_angle", "f64"}});
        S("Output2", {{"acceleration", "S:Axes"}, {"dangular_momentum_dt", "S:Axes"}});
        S("Coordinates", {{"position", "S:Axes"}, {"velocity", "S:Axes"}});
        // rotational flattening (rotational_flattening/common.rs:8-91)
        S("RotationalFlattening", {{"effect", "E:RotationalFlatteningEffect"}, {"parameters", "S:FlatteningParameters"},
                                   {"coordinates", "S:Coordinates"}});
        E("RotationalFlatteningEffect", {{"CentralBody", "E:RotationalFlatteningModel"},
                                         {"OrbitingBody", "E:RotationalFlatteningModel"}, {"Disabled", ""}});
        E("RotationalFlatteningModel", {{"OblateSpheroid", "S:OblateSpheroidParameters"}, {"CreepCoplanar", "!"}});
        S("OblateSpheroidParameters", {{"love_number", "f64"}});
        S("FlatteningParameters", {{"internal", "S:FlatteningInternal"}, {"output", "S:Output2"}});
        S("FlatteningInternal", {{"distance", "f64"}, {"scalar_product_of_vector_position_with_stellar_spin", "f64"},
                                 {"scalar_product_of_vector_position_with_planetary_spin", "f64"},
                                 {"radial_component_of_the_force_induced_by_rotation", "f64"},
                                 {"factor_for_the_force_induced_by_star_rotation", "f64"},
                                 {"factor_for_the_force_induced_by_planet_rotation", "f64"},
                                 {"orthogonal_component_of_the_force_induced_by_star_rotation", "f64"},
                                 {"orthogonal_component_of_the_force_induced_by_planet_rotation", "f64"}, {"shape", "S:Axes"}});
        // general relativity (general_relativity.rs:10-84)
        S("GeneralRelativity", {{"effect", "E:GeneralRelativityEffect"}, {"parameters", "S:GRParameters"},
                                {"coordinates", "S:Coordinates"}});
        E("GeneralRelativityEffect", {{"CentralBody", "E:GeneralRelativityImplementation"}, {"OrbitingBody", ""}, {"Disabled", ""}});
        E("GeneralRelativityImplementation", {{"Kidder1995", ""}, {"Anderson1975", ""}, {"Newhall1983", ""}, {"Disabled", ""}});
        S("GRParameters", {{"internal", "S:GRInternal"}, {"output", "S:Output2"}});
        S("GRInternal", {{"distance", "f64"}, {"radial_velocity", "f64"}, {"norm_velocity_vector", "f64"},
                         {"norm_velocity_vector_2", "f64"}, {"factor", "f64"}});
        // wind (wind.rs:8-40)
        S("Wind", {{"effect", "E:WindEffect"}, {"parameters", "S:WindParameters"}});
        E("WindEffect", {{"Interaction", ""}, {"Disabled", ""}});
        S("WindParameters", {{"input", "S:WindInput"}, {"internal", "S:WindInternal"}, {"output", "S:WindOutput"}});
        S("WindInput", {{"k_factor", "f64"}, {"rotation_saturation", "f64"}});
        S("WindInternal", {{"rotation_saturation_2", "f64"}});
        S("WindOutput", {{"dangular_momentum_dt", "S:Axes"}});
        // disk (disk.rs:7-60)
        S("Disk", {{"effect", "E:DiskEffect"}, {"parameters", "S:DiskParameters"}, {"coordinates", "S:Coordinates"}});
        E("DiskEffect", {{"CentralBody", "S:DiskProperties"}, {"OrbitingBody", ""}, {"Disabled", ""}});
        S("DiskProperties", {{"inner_edge_distance", "f64"}, {"outer_edge_distance", "f64"}, {"lifetime", "f64"},
                             {"alpha", "f64"}, {"surface_density_normalization", "f64"}, {"mean_molecular_weight", "f64"}});
        S("DiskParameters", {{"internal", "S:DiskInternal"}, {"output", "S:DiskOutput"}});
        S("DiskInternal", {{"distance", "f64"}, {"norm_velocity_vector", "f64"}, {"norm_velocity_vector_2", "f64"},
                           {"migration_timescale", "f64"}});
        S("DiskOutput", {{"acceleration", "S:Axes"}});
        // evolution (evolution.rs:9-40)
        E("EvolutionType", {{"GalletBolmont2017", "f64"}, {"BolmontMathis2016", "f64"}, {"Baraffe2015", "f64"},
                            {"Leconte2011", "f64"}, {"Baraffe1998", "f64"}, {"LeconteChabrier2013", "bool"}, {"NonEvolving", ""}});
        S("Evolver", {{"evolution", "E:EvolutionType"}, {"time", "V:f64"}, {"radius", "V:f64"},
                      {"radius_of_gyration_2", "V:f64"}, {"love_number", "V:f64"}, {"inverse_tidal_q_factor", "V:f64"},
                      {"left_index", "u64"}});
    }
};

inline const Schema &schema() {
    static const Schema s;
    return s;
}

// The value a field takes when a document leaves it out: what the
// reference's Python writer puts there for an unused field or a dummy particle
// (posidonius/particles/particle.py, effects/*.py): zeros, false, empty
// collections, and for an enum its "Disabled" variant, else its first unit
// variant (NonEvolving, MostMassiveParticle, ...).
inline Value default_value(const std::string &type) {
    const Schema &S = schema();
    Value v;
    if (type == "f64") return Value::number(0.);
    if (type == "u64") return Value::integer_value(0);
    if (type == "bool") { v.kind = Value::Bool; return v; }
    if (type.compare(0, 2, "S:") == 0) {
        v.kind = Value::Object;
        for (const auto &f : S.structs.at(type.substr(2))) v.members.emplace_back(f.first, default_value(f.second));
        return v;
    }
    if (type.compare(0, 2, "E:") == 0) {
        const auto &variants = S.enums.at(type.substr(2));
        const std::pair<std::string, std::string> *pick = nullptr;
        for (const auto &var : variants)
            if (var.second.empty() && (var.first == "Disabled" || !pick)) pick = &var;
        if (!pick) throw std::runtime_error("schema: enum " + type + " has no unit variant");
        v.kind = Value::String;
        v.text = pick->first;
        return v;
    }
    if (type[0] == 'A') {
        const size_t colon = type.find(':');
        const size_t n = std::stoul(type.substr(1, colon - 1));
        v.kind = Value::Array;
        for (size_t i = 0; i < n; ++i) v.items.push_back(default_value(type.substr(colon + 1)));
        return v;
    }
    if (type.compare(0, 2, "V:") == 0) { v.kind = Value::Array; return v; }
    if (type == "M") { v.kind = Value::Object; return v; }
    throw std::runtime_error("schema: unknown type " + type);
}

// ---- encoding: document -> bytes -------------------------------------------
// Fields a document leaves out are encoded as default_value(); fixed arrays
// given shorter are padded with it (a case lists only its n_particles real
// particles and evolvers when it is not written by the reference), and the
// evolver list is padded to MAX_PARTICLES = 10 entries, as Universe::new
// builds it (universe.rs:114-127).
class Encoder {
  public:
    std::string out;

    void value(const Value &v, const std::string &type, const std::string &where) {
        const Schema &S = schema();
        if (type == "f64") {
            const double x = (v.kind == Value::Null) ? std::nan("") : v.num();   // serde_json writes non-finite as null
            put(x);
        } else if (type == "u64") {
            if (!v.is_number()) throw std::runtime_error(where + ": expected an integer");
            put((uint64_t)std::strtoull(v.text.c_str(), nullptr, 10));
        } else if (type == "bool") {
            if (v.kind != Value::Bool) throw std::runtime_error(where + ": expected a bool");
            out.push_back(v.b ? 1 : 0);
        } else if (type.compare(0, 2, "S:") == 0) {
            const auto &fields = S.structs.at(type.substr(2));
            if (!v.is_object()) throw std::runtime_error(where + ": expected an object");
            for (const auto &f : fields) {
                const Value *m = v.find(f.first);
                if (m) {
                    value(*m, f.second, where + "." + f.first);
                } else {
                    value(default_value(f.second), f.second, where + "." + f.first);
                }
            }
        } else if (type.compare(0, 2, "E:") == 0) {
            const auto &variants = S.enums.at(type.substr(2));
            std::string tag;
            const Value *payload = nullptr;
            if (v.is_string()) {
                tag = v.str();
            } else if (v.is_object() && v.members.size() == 1) {
                tag = v.members[0].first;
                payload = &v.members[0].second;
            } else {
                throw std::runtime_error(where + ": expected an enum value");
            }
            for (size_t i = 0; i < variants.size(); ++i) {
                if (variants[i].first != tag) continue;
                if (variants[i].second == "!") throw std::runtime_error(where + ": variant " + tag + " is not supported");
                put((uint32_t)i);
                if (!variants[i].second.empty()) {
                    if (!payload) throw std::runtime_error(where + ": variant " + tag + " needs a payload");
                    value(*payload, variants[i].second, where + "." + tag);
                }
                return;
            }
            throw std::runtime_error(where + ": unknown variant " + tag);
        } else if (type[0] == 'A') {
            const size_t colon = type.find(':');
            const size_t n = std::stoul(type.substr(1, colon - 1));
            if (!v.is_array() || v.items.size() > n)
                throw std::runtime_error(where + ": more than " + std::to_string(n) + " elements do not fit the "
                                         "reference's layout (MAX_PARTICLES = 10)");
            const std::string et = type.substr(colon + 1);
            for (size_t i = 0; i < n; ++i)
                value(i < v.items.size() ? v.items[i] : default_value(et), et, where + "[" + std::to_string(i) + "]");
        } else if (type.compare(0, 2, "V:") == 0) {
            if (!v.is_array()) throw std::runtime_error(where + ": expected an array");
            const std::string et = type.substr(2);
            const size_t n = (et == "S:Evolver") ? std::max<size_t>(v.items.size(), 10) : v.items.size();
            put((uint64_t)n);
            for (size_t i = 0; i < n; ++i)
                value(i < v.items.size() ? v.items[i] : default_value(et), et, where + "[" + std::to_string(i) + "]");
        } else if (type == "M") {
            if (!v.is_object()) throw std::runtime_error(where + ": expected a map");
            std::map<uint64_t, double> sorted;   // bincode keeps the map's iteration order; any order reads back
            for (const auto &m : v.members) sorted[std::stoull(m.first)] = m.second.num();
            put((uint64_t)sorted.size());
            for (const auto &kv : sorted) { put(kv.first); put(kv.second); }
        } else {
            throw std::runtime_error("schema: unknown type " + type);
        }
    }

  private:
    template <class T> void put(T x) {
        char b[sizeof(T)];
        std::memcpy(b, &x, sizeof(T));   // the host is little endian, as bincode's default
        out.append(b, sizeof(T));
    }
};

inline std::string encode(const Value &doc) {
    Encoder e;
    e.value(doc, "S:WHFast", "WHFast");
    return std::move(e.out);
}

// ---- decoding: bytes -> document -------------------------------------------
class Decoder {
  public:
    Decoder(const std::string &bytes) : b_(bytes) {}

    Value value(const std::string &type, const std::string &where) {
        const Schema &S = schema();
        Value v;
        if (type == "f64") {
            const double x = get<double>(where);
            if (std::isfinite(x)) v = Value::number(x);
            // non-finite: null, as serde_json writes it
        } else if (type == "u64") {
            v.kind = Value::Number;
            v.text = std::to_string(get<uint64_t>(where));
        } else if (type == "bool") {
            const uint8_t c = get<uint8_t>(where);
            if (c > 1) throw std::runtime_error(where + ": invalid bool");
            v.kind = Value::Bool;
            v.b = c == 1;
        } else if (type.compare(0, 2, "S:") == 0) {
            v.kind = Value::Object;
            for (const auto &f : S.structs.at(type.substr(2)))
                v.members.emplace_back(f.first, value(f.second, where + "." + f.first));
        } else if (type.compare(0, 2, "E:") == 0) {
            const auto &variants = S.enums.at(type.substr(2));
            const uint32_t i = get<uint32_t>(where);
            if (i >= variants.size()) throw std::runtime_error(where + ": invalid variant index");
            if (variants[i].second == "!") throw std::runtime_error(where + ": variant " + variants[i].first + " is not supported");
            if (variants[i].second.empty()) {
                v.kind = Value::String;
                v.text = variants[i].first;
            } else {
                v.kind = Value::Object;
                v.members.emplace_back(variants[i].first, value(variants[i].second, where + "." + variants[i].first));
            }
        } else if (type[0] == 'A') {
            const size_t colon = type.find(':');
            const size_t n = std::stoul(type.substr(1, colon - 1));
            v.kind = Value::Array;
            for (size_t i = 0; i < n; ++i) v.items.push_back(value(type.substr(colon + 1), where + "[" + std::to_string(i) + "]"));
        } else if (type.compare(0, 2, "V:") == 0) {
            const uint64_t n = get<uint64_t>(where);
            if (n > remaining() / 8 + 1) throw std::runtime_error(where + ": length beyond the data");
            v.kind = Value::Array;
            for (uint64_t i = 0; i < n; ++i) v.items.push_back(value(type.substr(2), where + "[" + std::to_string(i) + "]"));
        } else if (type == "M") {
            const uint64_t n = get<uint64_t>(where);
            if (n > remaining() / 16) throw std::runtime_error(where + ": length beyond the data");
            v.kind = Value::Object;
            for (uint64_t i = 0; i < n; ++i) {
                const uint64_t key = get<uint64_t>(where);
                const double x = get<double>(where);
                v.members.emplace_back(std::to_string(key), std::isfinite(x) ? Value::number(x) : Value());
            }
        } else {
            throw std::runtime_error("schema: unknown type " + type);
        }
        return v;
    }

    size_t remaining() const { return b_.size() - i_; }

  private:
    const std::string &b_;
    size_t i_ = 0;

    template <class T> T get(const std::string &where) {
        if (remaining() < sizeof(T)) throw std::runtime_error(where + ": snapshot truncated");
        T x;
        std::memcpy(&x, b_.data() + i_, sizeof(T));
        i_ += sizeof(T);
        return x;
    }
};

// Bytes after the integrator are ignored, as bincode::deserialize_from does
// (output.rs:243-250).
inline Value decode(const std::string &bytes) {
    Decoder d(bytes);
    return d.value("S:WHFast", "WHFast");
}

}  // namespace posbin
