// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C entry point around the reference's own run_sanity_check
// (src/sanity_check.cpp:10-73, declared in src/sanity_check.h:5-6), which this
// recipe compiles from the reference's source where it lies.  The reference
// prints its count instead of returning it (sanity_check.cpp:65), so the harness
// captures stdout for the duration of the call and parses that line.
#include "sanity_check.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

extern "C" long long ref_sanity_count(const int* r, long long R,
                                      const int* s, long long S,
                                      double* seconds)
{
    std::fflush(stdout);
    FILE* cap = std::tmpfile();
    if (!cap) return -1;
    int saved = dup(STDOUT_FILENO);
    dup2(fileno(cap), STDOUT_FILENO);
    auto t0 = std::chrono::steady_clock::now();
    run_sanity_check(r, R, s, S);
    auto t1 = std::chrono::steady_clock::now();
    std::fflush(stdout);
    dup2(saved, STDOUT_FILENO);
    close(saved);
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();

    long long count = -1;
    std::rewind(cap);
    char line[512];
    const char* tag = "[sanityCheck] Answer Num (hash-based): ";
    while (std::fgets(line, sizeof(line), cap)) {
        const char* p = std::strstr(line, tag);
        if (p) count = std::atoll(p + std::strlen(tag));
    }
    std::fclose(cap);
    return count;
}
