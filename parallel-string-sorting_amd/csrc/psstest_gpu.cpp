// psstest_gpu.cpp -- runner with the reference psstest's selection, timing,
// checking and RESULT-line conventions (src/psstest.cpp), for the MI355X
// contestants registered below.
//
//   psstest_gpu [-a substr]... [-A name]... [-s size] [-r repeats] [-N] [-L] [--suffix] input
//
// input: random2 | random4 | random10 | random62 | random255 (the reference's
// generators, src/tools/input.hpp:425-468), ascii (length U[1,64] over bytes
// 33..126), or a file whose lines become the strings (load_plain,
// src/tools/input.hpp:225-300), possibly compressed with its decompressed size
// in the name, e.g. text.1000000.gz (load_compressed, input.hpp:303-422).
// --suffix sorts every suffix of the file instead (src/psstest.cpp:502-518).
// -s limits the input size in bytes (K/M/G suffixes).  Each run times the sort call only (CLOCK_MONOTONIC wall clock,
// src/psstest.cpp:593-603), then checks order, permutation and (for LCP
// contestants) every LCP value and (for LCP-cache contestants) every
// distinguishing character, as verify_lcp_cache (src/tools/stringptr.hpp:977-1031).
#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <csignal>
#include <getopt.h>
#include <sys/wait.h>
#include <unistd.h>
#include <vector>

#include "../../include/ps5_mi355x.h"
#include "contest.hpp"

namespace pss {
Contest* getContestSingleton()
{
    static Contest* c = new Contest;
    return c;
}
}  // namespace pss

// ---- contestants -----------------------------------------------------------


// LCP merge sort in the shape of the reference's eberle/parallel_lcp_mergesort
// contestants (src/parallel/eberle-parallel-lcp-mergesort.hpp:99-126): eight
// index ranges sorted with their LCP arrays, then merged by the LCP-aware
// device merge -- the path a sharded (multi-GPU) sort takes.
static void mi355x_lcp_mergesort_8way(unsigned char** strings, size_t n)
{
    const unsigned K = n >= 8 ? 8 : 1;
    std::vector<unsigned char*> runs_v[8];
    std::vector<uintptr_t> lcps_v[8];
    unsigned char* const* runs[8];
    const uintptr_t* lcps[8];
    size_t sizes[8];
    for (unsigned r = 0; r < K; ++r) {
        const size_t b = n * r / K, e = n * (r + 1) / K;
        runs_v[r].assign(strings + b, strings + e);
        lcps_v[r].assign(e - b + 1, 0);
        mi355x_ps5_sort_lcp(runs_v[r].data(), lcps_v[r].data(), e - b);
        runs[r] = runs_v[r].data();
        lcps[r] = lcps_v[r].data();
        sizes[r] = e - b;
    }
    if (int rc = ps5_lcp_merge_strings(runs, lcps, sizes, K, strings, nullptr)) {
        fprintf(stderr, "ps5_lcp_merge_strings: error %d: %s\n", rc, ps5_last_error());
        abort();
    }
}

static void std_sort_strcmp(unsigned char** strings, size_t n)
{
    std::stable_sort(strings, strings + n, [](const unsigned char* a, const unsigned char* b) {
        return strcmp((const char*)a, (const char*)b) < 0;
    });
}

PSS_CONTESTANT_PARALLEL(mi355x_ps5_sort, "mi355x/parallel_sample_sortBTCU2",
                        "pS5 on MI355X: splitter tree, key cache, LDS blocks (gfx950)")
PSS_CONTESTANT_PARALLEL(mi355x_ps5_sort_lcp, "mi355x/parallel_sample_sortBTCU2_lcp",
                        "pS5 on MI355X with LCP array output (gfx950)")
PSS_CONTESTANT_PARALLEL(mi355x_ps5_sort_lcp_cache, "mi355x/parallel_sample_sortBTCU2_lcp_cache",
                        "pS5 on MI355X with LCP array and distinguishing characters (gfx950)")
PSS_CONTESTANT_PARALLEL(mi355x_ps5_sort_out, "mi355x/parallel_sample_sortBTCU2_out",
                        "pS5 on MI355X, separate output array (gfx950)")
PSS_CONTESTANT_PARALLEL(mi355x_lcp_mergesort_8way, "mi355x/lcp_mergesort_8way",
                        "8 sorted index ranges merged by the device LCP merge (gfx950)")
PSS_CONTESTANT(std_sort_strcmp, "std/stable_sort_strcmp", "std::stable_sort with strcmp (host, one thread)")

// ---- input -----------------------------------------------------------------

struct Input {
    std::vector<uint8_t> data;   // NUL-separated strings + 16 padding NULs
    std::vector<uint32_t> offs;
    std::string name;
};

static uint64_t parse_size(const char* s)
{
    char* e = nullptr;
    double v = strtod(s, &e);
    uint64_t m = 1;
    if (e && *e) {
        switch (*e) {
        case 'k': case 'K': m = 1ull << 10; break;
        case 'm': case 'M': m = 1ull << 20; break;
        case 'g': case 'G': m = 1ull << 30; break;
        default: break;
        }
    }
    return (uint64_t)(v * (double)m);
}

static bool g_suffix = false;   // --suffix: sort all suffixes of the input file

// Whole file, at most `limit` bytes (load_plain, src/tools/input.hpp:225-300);
// data gets 16 zero bytes of padding.
static bool read_plain(const std::string& path, uint64_t limit, std::vector<uint8_t>& data, uint64_t& fs)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) { fprintf(stderr, "cannot open %s: %s\n", path.c_str(), strerror(errno)); return false; }
    fseek(f, 0, SEEK_END);
    fs = (uint64_t)ftell(f);
    rewind(f);
    if (limit && fs > limit) fs = limit;
    data.assign(fs + 16, 0);
    const bool ok = fread(data.data(), 1, fs, f) == fs;
    fclose(f);
    if (!ok) fprintf(stderr, "cannot read %s\n", path.c_str());
    return ok;
}

// .gz/.bz2/.xz/.lzo/.zst files whose name carries the decompressed size
// before the extension ("name.<bytes>.gz"), read through the decompressor
// (load_compressed, src/tools/input.hpp:303-422).  handled = false when the
// name has no such extension.
static bool read_compressed(const std::string& path, uint64_t limit, std::vector<uint8_t>& data, uint64_t& fs,
                            bool& handled)
{
    static const struct { const char* ext; const char* prog; } tools[] = {
        { ".gz", "gzip" }, { ".bz2", "bzip2" }, { ".xz", "xz" }, { ".lzo", "lzop" }, { ".zst", "zstd" },
    };
    const char* prog = nullptr;
    for (auto& t : tools) {
        const size_t el = strlen(t.ext);
        if (path.size() > el && path.compare(path.size() - el, el, t.ext) == 0) prog = t.prog;
    }
    handled = prog != nullptr;
    if (!prog) return true;
    const std::string::size_type dot = path.rfind('.');
    std::string::size_type i = dot;
    uint64_t size = 0, v = 1;
    while (i > 0 && isdigit((unsigned char)path[i - 1])) {
        size += (uint64_t)(path[i - 1] - '0') * v;
        v *= 10;
        --i;
    }
    if (size == 0 || i == 0 || path[i - 1] != '.') {
        fprintf(stderr, "could not find the decompressed size in the file name %s\n", path.c_str());
        return false;
    }
    if (limit && size > limit) size = limit;
    int fd[2];
    if (pipe(fd) != 0) { fprintf(stderr, "pipe: %s\n", strerror(errno)); return false; }
    const pid_t pid = fork();
    if (pid < 0) { fprintf(stderr, "fork: %s\n", strerror(errno)); return false; }
    if (pid == 0) {     // child: the decompressor writes into the pipe
        close(fd[0]);
        dup2(fd[1], STDOUT_FILENO);
        execlp(prog, prog, "-dc", path.c_str(), (char*)nullptr);
        _exit(127);
    }
    close(fd[1]);
    data.assign(size + 16, 0);
    uint64_t pos = 0;
    while (pos < size) {
        const ssize_t rb = read(fd[0], data.data() + pos, std::min<uint64_t>(8u << 20, size - pos));
        if (rb <= 0) break;
        pos += (uint64_t)rb;
    }
    close(fd[0]);
    kill(pid, SIGTERM);
    int status = 0;
    waitpid(pid, &status, 0);
    if (pos < size) { fprintf(stderr, "%s: %s gave %" PRIu64 " of %" PRIu64 " bytes\n", path.c_str(), prog, pos, size); return false; }
    fs = size;
    return true;
}

static bool load_input(const std::string& path, uint64_t size, Input& in)
{
    static const struct { const char* name; const char* letters; } gens[] = {
        { "random2", "01" }, { "random4", "ACGT" }, { "random10", "0123456789" },
        { "random62", "0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz" },
    };
    in.name = path;
    for (auto& g : gens) {
        if (path == g.name) {
            if (!size) { fprintf(stderr, "random input size must be given with -s\n"); return false; }
            in.data.assign(size + 9 + 16, 0);
            in.offs.resize(size / 16 + 2);
            uint64_t n = ps5_gen_reference_random(size, (const uint8_t*)g.letters, (uint32_t)strlen(g.letters),
                                                  in.data.data(), in.offs.data(), in.offs.size());
            in.offs.resize(n);
            return true;
        }
    }
    if (path == "random255") {
        if (!size) { fprintf(stderr, "random input size must be given with -s\n"); return false; }
        uint8_t letters[255];
        for (int i = 0; i < 255; ++i) letters[i] = (uint8_t)(i + 1);
        in.data.assign(size + 9 + 16, 0);
        in.offs.resize(size / 16 + 2);
        uint64_t n = ps5_gen_reference_random(size, letters, 255, in.data.data(), in.offs.data(), in.offs.size());
        in.offs.resize(n);
        return true;
    }
    if (path == "ascii") {
        if (!size) size = 2ull << 30;
        uint8_t alpha[94];
        for (int i = 0; i < 94; ++i) alpha[i] = (uint8_t)(33 + i);
        in.data.assign(size + 16, 0);
        in.offs.resize(size / 2 + 1);
        uint64_t used = 0;
        uint64_t n = ps5_gen_uniform(0x5EED0002ull, size, 1, 64, alpha, 94, in.data.data(), in.offs.data(),
                                     in.offs.size(), &used);
        in.offs.resize(n);
        return true;
    }
    uint64_t fs = 0;
    bool handled = false;
    if (!read_compressed(path, size, in.data, fs, handled)) return false;
    if (!handled && !read_plain(path, size, in.data, fs)) return false;
    if (fs == 0) { fprintf(stderr, "%s is empty\n", path.c_str()); return false; }
    if (fs > 0xFFFFFFF0ull) { fprintf(stderr, "%s: input above 4 GiB\n", path.c_str()); return false; }
    if (g_suffix) {   // every position starts a string (src/psstest.cpp:502-518)
        in.data[fs - 1] = 0;
        in.offs.resize(fs);
        for (uint64_t i = 0; i < fs; ++i) in.offs[i] = (uint32_t)i;
        in.name += " (suffixes)";
        return true;
    }
    for (uint64_t i = 0; i < fs; ++i)
        if (in.data[i] == '\n') in.data[i] = 0;
    in.data[fs - 1] = 0;
    for (uint64_t i = 0; i < fs; ++i)
        if (i == 0 || in.data[i - 1] == 0) in.offs.push_back((uint32_t)i);
    return true;
}

// ---- checking --------------------------------------------------------------

static bool check(const Input& in, unsigned char** s, const uintptr_t* lcp, const unsigned char* cache, size_t n,
                  std::string& why)
{
    // permutation: every input string exactly once (by address)
    std::vector<const unsigned char*> a(s, s + n), b(n);
    for (size_t i = 0; i < n; ++i) b[i] = in.data.data() + in.offs[i];
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    if (a != b) { why = "not a permutation"; return false; }
    for (size_t i = 1; i < n; ++i) {
        const unsigned char *x = s[i - 1], *y = s[i];
        size_t h = 0;
        while (x[h] && x[h] == y[h]) ++h;
        if (x[h] > y[h]) { why = "invalid order at " + std::to_string(i); return false; }
        if (lcp && lcp[i] != h) { why = "wrong lcp at " + std::to_string(i); return false; }
        if (cache && cache[i] != y[h]) { why = "wrong cache at " + std::to_string(i); return false; }
    }
    return true;
}

int main(int argc, char** argv)
{
    std::vector<std::string> sel, sel_full;
    uint64_t size = 0;
    int repeats = 1, opt;
    bool no_check = false, list = false;
    static const option longopts[] = { { "suffix", no_argument, nullptr, 'S' }, { nullptr, 0, nullptr, 0 } };
    while ((opt = getopt_long(argc, argv, "a:A:s:r:NLS", longopts, nullptr)) != -1) {
        switch (opt) {
        case 'S': g_suffix = true; break;
        case 'a': sel.push_back(optarg); break;
        case 'A': sel_full.push_back(optarg); break;
        case 's': size = parse_size(optarg); break;
        case 'r': repeats = std::max(1, atoi(optarg)); break;
        case 'N': no_check = true; break;
        case 'L': list = true; break;
        default:
            fprintf(stderr, "usage: %s [-a substr] [-A name] [-s size] [-r repeats] [-N] [-L] [--suffix] input\n",
                    argv[0]);
            return 1;
        }
    }
    auto& cl = pss::getContestSingleton()->m_list;
    std::sort(cl.begin(), cl.end(),
              [](const pss::Contestant* x, const pss::Contestant* y) { return strcmp(x->m_algoname, y->m_algoname) < 0; });
    if (list) {
        for (auto* c : cl) printf("%-45s %s\n", c->m_algoname, c->m_description);
        return 0;
    }
    if (optind >= argc) { fprintf(stderr, "no input given\n"); return 1; }
    Input in;
    if (!load_input(argv[optind], size, in)) return 1;
    const size_t n = in.offs.size();
    const uint64_t chars = in.data.size() - 16;
    char host[128] = "";
    gethostname(host, sizeof host);
    int failures = 0;
    for (auto* c : cl) {
        bool take = sel.empty() && sel_full.empty();
        for (auto& s : sel) take |= strstr(c->m_algoname, s.c_str()) != nullptr;
        for (auto& s : sel_full) take |= strcmp(c->m_algoname, s.c_str()) == 0;
        if (!take) continue;
        for (int r = 0; r < repeats; ++r) {
            std::vector<unsigned char*> strings(n);
            for (size_t i = 0; i < n; ++i) strings[i] = in.data.data() + in.offs[i];
            const bool with_lcp = c->is_lcp_func() || c->is_lcp_cache_func();
            std::vector<uintptr_t> lcp;
            std::vector<unsigned char> cache;
            if (with_lcp) {
                lcp.assign(n, (uintptr_t)-1);
                if (n) lcp[0] = 42;   // must stay unchanged (src/psstest.cpp:578-584)
            }
            if (c->is_lcp_cache_func()) {   // src/psstest.cpp:586-591
                cache.assign(n, (unsigned char)-1);
                if (n) cache[0] = 0;
            }
            timespec t0, t1;
            clock_gettime(CLOCK_MONOTONIC, &t0);
            if (c->is_lcp_func()) c->m_run_lcp_func(strings.data(), lcp.data(), n);
            else if (c->is_lcp_cache_func()) c->m_run_lcp_cache_func(strings.data(), lcp.data(), cache.data(), n);
            else c->m_run_func(strings.data(), n);
            clock_gettime(CLOCK_MONOTONIC, &t1);
            const double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
            std::string status = "ok", why;
            if (!no_check) {
                if (!check(in, strings.data(), with_lcp ? lcp.data() : nullptr,
                           c->is_lcp_cache_func() ? cache.data() : nullptr, n, why))
                    status = "failed";
                if (with_lcp && n && lcp[0] != 42) { status = "failed"; why = "lcp[0] was written"; }
            } else {
                status = "unchecked";
            }
            char dt[64];
            time_t now = time(nullptr);
            strftime(dt, sizeof dt, "%Y-%m-%d %H:%M:%S", localtime(&now));
            printf("RESULT\tdatetime=%s\thost=%s\talgo=%s\tdata=%s\tchar_count=%" PRIu64 "\tstring_count=%zu"
                   "\ttime=%.6f\tchars_per_s=%.0f\tstatus=%s%s%s\n",
                   dt, host, c->m_algoname, in.name.c_str(), chars, n, secs, secs > 0 ? chars / secs : 0.0,
                   status.c_str(), why.empty() ? "" : "\tfailure=", why.c_str());
            fflush(stdout);
            if (status == "failed") failures++;
        }
    }
    return failures ? 2 : 0;
}
