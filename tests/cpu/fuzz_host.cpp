// fuzz_host.cpp -- bounded mutation fuzzing of the command-line host's input
// handling, built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/cpu/Makefile).  Each seed document (a case, JSON) is mutated by a
// fixed-seed generator: byte flips, deletions, insertions, duplicated spans
// and truncations.  Every mutant goes through the JSON parser and the case
// reader (Case: the document to the C ABI's arrays), and through the bincode
// snapshot decoder as bytes; every seed also goes JSON -> bincode -> document.
// Errors must surface as exceptions; a memory error or undefined behaviour
// aborts the process under the sanitizers.
//
// Usage: fuzz_host <iterations> <seed.json>...
#define main posidonius_cli_main
#include "../../posidonius_amd/host/posidonius_cli.cpp"
#undef main

#include <random>

// The command line's device calls are never made here; these stand-ins only
// satisfy the linker.
extern "C" {
int32_t pos_abi_version(void) { return POS_ABI_VERSION; }
const char *pos_last_error(void) { return "no device in the fuzzer"; }
int32_t pos_create(const pos_ensemble_desc *, const pos_options *, pos_ens **) { return POS_E_NO_DEVICE; }
int32_t pos_step(pos_ens *, int64_t) { return POS_E_NO_DEVICE; }
int32_t pos_get_state(pos_ens *, pos_state *) { return POS_E_NO_DEVICE; }
int32_t pos_history_enable(pos_ens *, int32_t) { return POS_E_NO_DEVICE; }
int32_t pos_history_drain(pos_ens *, pos_history_record *, int64_t, int64_t *) { return POS_E_NO_DEVICE; }
void pos_destroy(pos_ens *) {}
}

namespace {

std::string mutate(const std::string &s, std::mt19937_64 &rng) {
    std::string m = s;
    const int edits = 1 + (int)(rng() % 8);
    for (int e = 0; e < edits && !m.empty(); ++e) {
        const size_t i = rng() % m.size();
        switch (rng() % 6) {
            case 0: m[i] = (char)(rng() & 0xff); break;
            case 1: m.erase(i, 1 + rng() % 16); break;
            case 2: m.insert(i, 1, "{}[]\",:0123456789.eE-+ntrufals\\"[rng() % 31]); break;
            case 3: { const size_t n = 1 + rng() % 64; m.insert(i, m.substr(i, n)); break; }
            case 4: m.resize(i); break;
            default: m[i] = "{}[]\",:0-9"[rng() % 10]; break;
        }
    }
    return m;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) { std::fprintf(stderr, "usage: fuzz_host <iterations> <seed.json>...\n"); return 2; }
    const long iterations = std::atol(argv[1]);
    std::mt19937_64 rng(20261203);
    long parsed = 0, cases = 0, decoded = 0;
    for (int a = 2; a < argc; ++a) {
        const std::string text = read_file(argv[a]);
        const Value doc = posjson::parse(text);
        const std::string bin = posbin::encode(doc);
        const Value back = posbin::decode(bin);
        if (posbin::encode(back) != bin) { std::fprintf(stderr, "bincode round trip differs for %s\n", argv[a]); return 1; }
        for (long i = 0; i < iterations; ++i) {
            const std::string t = mutate(text, rng);
            try {
                const Value v = posjson::parse(t);
                ++parsed;
                Case c(v);
                const pos_ensemble_desc d = c.desc();
                (void)d;
                ++cases;
                (void)posbin::encode(v);
            } catch (const std::exception &) {
            }
            try {
                (void)posbin::decode(mutate(bin, rng));
                ++decoded;
            } catch (const std::exception &) {
            }
        }
    }
    std::printf("fuzz_host: %ld JSON mutants parsed, %ld read as cases, %ld bincode mutants decoded\n", parsed, cases,
                decoded);
    return 0;
}
