/*
 * mcref_rng.c - deterministic replacement for the Fortran RANDOM_NUMBER /
 * RANDOM_SEED intrinsics, linked ONLY into the reference harness
 * (oracle/_ref/mcref_harness, built from the unmodified reference sources).
 *
 * The reference seeds random_number from the wall clock
 * (photon_mod.f90:68-85), so its runs cannot be replayed.  Linking these
 * definitions ahead of the Fortran runtime makes every random_number call
 * return the next uniform of a Philox4x32-10 stream selected by
 * mcref_set_stream(seed, stream, packet): the same stream the oracle and the
 * GPU kernels draw for that packet, u = (word >> 8) * 2^-24.
 *
 * Entry points and argument types follow flang's public runtime header
 * include/flang/Runtime/random.h; descriptors are CFI_cdesc_t
 * (ISO_Fortran_binding.h).
 */
#include <stdint.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>
#include "ISO_Fortran_binding.h"

static uint64_t g_seed = 0, g_packet = 0;
static uint32_t g_stream = 0, g_k = 0;
static uint64_t g_total = 0;

static void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static float next_uniform(void)
{
    uint32_t ctr[4] = { g_k >> 2, (uint32_t)g_packet, (uint32_t)(g_packet >> 32), g_stream };
    uint32_t key[2] = { (uint32_t)g_seed, (uint32_t)(g_seed >> 32) };
    uint32_t o[4];
    philox(ctr, key, o);
    float u = (float)(o[g_k & 3] >> 8) * 0x1p-24f;
    g_k++;
    g_total++;
    return u;
}

/* called from the Fortran harness through bind(C) */
void mcref_set_stream(const int64_t *seed, const int32_t *stream, const int64_t *packet)
{
    g_seed = (uint64_t)*seed;
    g_stream = (uint32_t)*stream;
    g_packet = (uint64_t)*packet;
    g_k = 0;
}

void mcref_draw_count(int64_t *n) { *n = (int64_t)g_k; }

static size_t desc_count(const CFI_cdesc_t *d)
{
    size_t n = 1;
    for (int r = 0; r < d->rank; ++r) n *= (size_t)d->dim[r].extent;
    return n;
}

void _FortranARandomNumber(const CFI_cdesc_t *harvest, const char *source, int line)
{
    (void)source; (void)line;
    size_t n = desc_count(harvest);
    if (harvest->elem_len == 4) {
        float *p = (float *)harvest->base_addr;
        /* contiguous or scalar harvests only (all the reference uses) */
        for (size_t i = 0; i < n; ++i) p[i] = next_uniform();
    } else if (harvest->elem_len == 8) {
        double *p = (double *)harvest->base_addr;
        for (size_t i = 0; i < n; ++i) p[i] = (double)next_uniform();
    } else {
        fprintf(stderr, "mcref_rng: unsupported harvest kind %zu\n", harvest->elem_len);
        abort();
    }
}

void _FortranARandomSeedSize(const CFI_cdesc_t *size, const char *source, int line)
{
    (void)source; (void)line;
    if (!size || !size->base_addr) return;
    if (size->elem_len == 8) *(int64_t *)size->base_addr = 1;
    else *(int32_t *)size->base_addr = 1;
}

void _FortranARandomSeedGet(const CFI_cdesc_t *get, const char *source, int line)
{
    (void)source; (void)line;
    if (!get || !get->base_addr) return;
    memset(get->base_addr, 0, desc_count(get) * get->elem_len);
}

void _FortranARandomSeedPut(const CFI_cdesc_t *put, const char *source, int line)
{
    (void)put; (void)source; (void)line;
}
