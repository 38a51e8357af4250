/* Helpers shared by the converted workload programs, their host handlers and
 * their HIP handlers (so the same inline code runs on the host and the GPU). */
#ifndef DVMH_PROG_COMMON_H
#define DVMH_PROG_COMMON_H

#ifdef __HIPCC__
#define DVMH_HD __host__ __device__
#else
#define DVMH_HD
#endif

#define Max(a, b) ((a) > (b) ? (a) : (b))
#define Min(a, b) ((a) < (b) ? (a) : (b))

/* Seeded portable generator: SplitMix64 finaliser of seed + (n + 1) * golden,
 * top 53 bits scaled to [0, 1).  Pure integer arithmetic plus one exact scaling,
 * so host and GPU produce bit-identical values. */
static inline DVMH_HD double prog_u01(unsigned long long seed, unsigned long long n) {
    unsigned long long z = seed + (n + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static inline DVMH_HD unsigned long long prog_mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

#endif
