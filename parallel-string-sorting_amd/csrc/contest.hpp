// contest.hpp -- plug-in registry with the reference's Contest semantics.
//
// Mirrors src/tools/contest.hpp:38-186 of the reference: contestants register
// themselves at static-initialisation time with a name, a description and a
// function of one of the Contest signatures (func_type / func_lcp_type /
// func_lcp_cache_type, src/tools/contest.hpp:88-91), and
// the runner selects them by substring (-a) or exact name (-A), as
// gopt_algorithm_select does (src/psstest.cpp:171-199).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace pss {

class Contestant {
public:
    typedef void (*func_type)(unsigned char** strings, size_t n);
    typedef void (*func_lcp_type)(unsigned char** strings, uintptr_t* lcp, size_t n);
    typedef void (*func_lcp_cache_type)(unsigned char** strings, uintptr_t* lcp, unsigned char* cache, size_t n);

    const char* m_algoname;
    const char* m_description;
    func_type m_run_func = nullptr;
    func_lcp_type m_run_lcp_func = nullptr;
    func_lcp_cache_type m_run_lcp_cache_func = nullptr;
    bool m_parallel;

    Contestant(func_type f, const char* name, const char* desc, bool parallel);
    Contestant(func_lcp_type f, const char* name, const char* desc, bool parallel);
    Contestant(func_lcp_cache_type f, const char* name, const char* desc, bool parallel);

    bool is_lcp_func() const { return m_run_lcp_func != nullptr; }
    bool is_lcp_cache_func() const { return m_run_lcp_cache_func != nullptr; }
    bool is_parallel() const { return m_parallel; }
};

class Contest {
public:
    std::vector<Contestant*> m_list;
    void register_contestant(Contestant* c) { m_list.push_back(c); }
};

Contest* getContestSingleton();

inline Contestant::Contestant(func_type f, const char* name, const char* desc, bool parallel)
    : m_algoname(name), m_description(desc), m_run_func(f), m_parallel(parallel)
{
    getContestSingleton()->register_contestant(this);
}

inline Contestant::Contestant(func_lcp_type f, const char* name, const char* desc, bool parallel)
    : m_algoname(name), m_description(desc), m_run_lcp_func(f), m_parallel(parallel)
{
    getContestSingleton()->register_contestant(this);
}

inline Contestant::Contestant(func_lcp_cache_type f, const char* name, const char* desc, bool parallel)
    : m_algoname(name), m_description(desc), m_run_lcp_cache_func(f), m_parallel(parallel)
{
    getContestSingleton()->register_contestant(this);
}

}  // namespace pss

#define PSS_CONTESTANT(func, algoname, desc)                                      \
    static const pss::Contestant* _Contestant_##func##_register __attribute__((unused)) = \
        new pss::Contestant(func, algoname, desc, false);

#define PSS_CONTESTANT_PARALLEL(func, algoname, desc)                             \
    static const pss::Contestant* _Contestant_##func##_register __attribute__((unused)) = \
        new pss::Contestant(func, algoname, desc, true);
