/* Drives the CPU oracle's entry points on small scenes so that an AddressSanitizer /
 * UndefinedBehaviorSanitizer build (tests/test_sanitizers_cpu.py) checks its memory accesses:
 * list conversion of every vertex and colour type, the LBVH build (serial and threaded, 30- and
 * 63-bit keys), the treelet P-k-d build for several treelet sizes, and brute-force, LBVH and
 * P-k-d renders in every shading mode. Exit status 0 when every call succeeded and the three
 * renders agree. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t rng = 12345u;
static float urand(void) {
    rng = rng * 1664525u + 1013904223u;
    return (float)(rng >> 8) / 16777216.0f;
}

int main(void) {
    const uint64_t n = 3001;
    float* xyzr = malloc(n * 16);
    float* rgba = malloc(n * 16);
    for (uint64_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) xyzr[4 * i + a] = urand();
        xyzr[4 * i + 3] = 0.01f + 0.03f * urand();
        for (int a = 0; a < 4; ++a) rgba[4 * i + a] = urand();
    }
    memcpy(xyzr + 4 * 100, xyzr + 4 * 7, 16); /* coincident spheres */
    /* every list type through the ingest */
    double dv[30];
    uint16_t sv[30], us[40];
    uint8_t u8[40];
    float fv[40];
    for (int i = 0; i < 30; ++i) {
        dv[i] = urand();
        sv[i] = (uint16_t)(urand() * 60000.0f);
    }
    for (int i = 0; i < 40; ++i) {
        us[i] = (uint16_t)(urand() * 65535.0f);
        u8[i] = (uint8_t)(urand() * 255.0f);
        fv[i] = urand();
    }
    float tf[8] = {0, 0, 1, 1, 1, 1, 0, 1};
    int bad = 0;
    for (int vt = 1; vt <= 4; ++vt)
        for (int ct = 0; ct <= 7; ++ct) {
            orc_list l;
            memset(&l, 0, sizeof l);
            l.vert_type = vt;
            l.col_type = ct;
            l.vert = vt == 4 ? (const void*)dv : vt == 3 ? (const void*)sv : (const void*)fv;
            l.col = ct == 6 ? (const void*)us : (ct == 1 || ct == 2) ? (const void*)u8 : (const void*)dv;
            l.count = 5;
            l.global_radius = 0.1f;
            l.global_rgba[0] = 200;
            l.col_min = 0.0f;
            l.col_max = 1.0f;
            float out_x[20], out_c[20];
            uint64_t got = 0;
            if (orc_convert_lists(&l, 1, tf, 2, 1.0f, 0.0f, 1, out_x, out_c, &got) != 0 || got != 5) bad |= 1;
        }
    orc_camera cam;
    memset(&cam, 0, sizeof cam);
    cam.position[0] = 0.5f;
    cam.position[1] = 0.5f;
    cam.position[2] = 2.2f;
    cam.direction[2] = -1.0f;
    cam.up[1] = 1.0f;
    cam.fovy = 0.7f;
    cam.aspect = 4.0f / 3.0f;
    cam.near_plane = 0.1f;
    cam.far_plane = 10.0f;
    cam.tile_end[0] = cam.tile_end[1] = 1.0f;
    const uint32_t W = 40, H = 30;
    uint32_t* r8[3];
    float* col[3];
    int32_t* ids[3];
    uint16_t* aux[3];
    for (int k = 0; k < 3; ++k) {
        r8[k] = malloc(W * H * 4);
        col[k] = malloc(W * H * 16);
        ids[k] = malloc(W * H * 4);
        aux[k] = malloc(W * H * 16);
    }
    float* depth = malloc(W * H * 4);
    float* tt = malloc(W * H * 4);
    for (int bits = 30; bits <= 63; bits += 33)
        for (int threads = 1; threads <= 4; threads += 3) {
            orc_bvh b;
            if (orc_bvh_build_mt(xyzr, n, bits, threads, &b) != 0) {
                bad |= 2;
                continue;
            }
            for (uint32_t M = 2; M <= 1024; M *= 8) {
                orc_pkd k;
                if (orc_pkd_build(&b, M, threads, &k) != 0) {
                    bad |= 4;
                    continue;
                }
                for (int mode = 0; mode <= 2; ++mode) {
                    orc_params p;
                    memset(&p, 0, sizeof p);
                    p.width = W;
                    p.height = H;
                    p.spp = 2;
                    p.mode = mode;
                    p.ao_distance = 0.2f;
                    p.want_aux = 1;
                    p.max_bounces = mode == 2 ? 2 : 0;
                    uint64_t cnt[2];
                    orc_render_brute(xyzr, rgba, n, &cam, &p, r8[0], depth, col[0], tt, ids[0], aux[0]);
                    orc_render_bvh(&b, rgba, &cam, &p, threads, r8[1], depth, col[1], tt, ids[1], aux[1], cnt);
                    orc_render_pkd(&k, rgba, &cam, &p, threads, r8[2], depth, col[2], tt, ids[2], aux[2], cnt);
                    for (int q = 1; q < 3; ++q)
                        if (memcmp(ids[0], ids[q], W * H * 4) || memcmp(col[0], col[q], W * H * 16) ||
                            memcmp(aux[0], aux[q], W * H * 16))
                            bad |= 8;
                }
                orc_pkd_free(&k);
            }
            orc_bvh_free(&b);
        }
    for (int k = 0; k < 3; ++k) {
        free(r8[k]);
        free(col[k]);
        free(ids[k]);
        free(aux[k]);
    }
    free(depth);
    free(tt);
    free(xyzr);
    free(rgba);
    printf("oracle sanitize run: %s (flags %d)\n", bad ? "FAILED" : "ok", bad);
    return bad;
}
