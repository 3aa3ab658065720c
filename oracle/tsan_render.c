/*
 * tsan_render.c - drives the oracle's OpenMP render (orc_render) on a small
 * procedural scene with several threads, for a ThreadSanitizer build of the
 * oracle (oracle/Makefile, target tsan).  TEST INFRASTRUCTURE ONLY.
 *
 * Scene: a 24^3 grid holding a sphere and a floor slab of Lambert voxels
 * plus one metal voxel row and one emissive voxel (the NEE emissive branch),
 * pinhole and thin-lens cameras from orc_camera_matrices, the gradient
 * environment, 4 bounces.  Renders the frame twice with `threads` OpenMP
 * threads (FP64 sums on) and once with one thread, and checks that the
 * images, sums and counters are identical (the parallel loop writes
 * disjoint pixels and reduces its counters).  Prints "tsan_render ok" and a
 * checksum; exits 1 on a mismatch.
 */
#include "vt_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define N 24
#define W 40
#define H 28

static int render(const orc_scene* s, const orc_env* e, const orc_camera* c, const orc_settings* st, int threads,
                  float* acc, double* sum, orc_counters* cnt)
{
    memset(acc, 0, sizeof(float) * W * H * 4);
    memset(sum, 0, sizeof(double) * W * H * 3);
    memset(cnt, 0, sizeof(*cnt));
    orc_render(s, e, c, st, W, H, 0, H, 12345u, 0, 3, 0xffffffffu, acc, sum, cnt, threads);
    return 0;
}

int main(int argc, char** argv)
{
    const int threads = argc > 1 ? atoi(argv[1]) : 4;
    static int32_t grid[N * N * N];
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const float dx = x - 12.0f, dy = y - 13.0f, dz = z - 11.0f;
                int32_t off = -1;
                if (y < 3 || dx * dx + dy * dy + dz * dz < 36.0f) off = 8 * ((x + y + z) % 3);
                if (y == 3 && z == 4 && x > 4 && x < 18) off = 24;   /* metal row */
                grid[x + N * (y + N * z)] = off;
            }
    grid[20 + N * (3 + N * 20)] = 32;                                  /* emissive voxel */
    /* records: type, emission[3], albedo/reflectance[3], roughness */
    static const float mat[40] = {
        0, 0, 0, 0, 0.7f, 0.6f, 0.5f, 0,  0, 0, 0, 0, 0.3f, 0.6f, 0.4f, 0,  0, 0, 0, 0, 0.5f, 0.5f, 0.8f, 0,
        1, 0, 0, 0, 0.9f, 0.8f, 0.7f, 12.0f,  0, 4.0f, 3.0f, 2.0f, 0.5f, 0.5f, 0.5f, 0 };
    int32_t emissive[1] = { 20 + N * (3 + N * 20) };
    orc_scene s;
    memset(&s, 0, sizeof(s));
    s.res[0] = s.res[1] = s.res[2] = N;
    for (int k = 0; k < 3; ++k) {
        s.bmin[k] = -500.0f;
        s.bmax[k] = 500.0f;
        s.vsize[k] = 1000.0f / N;
    }
    s.offsets = grid;
    s.mat = mat;
    s.mat_size = 40;
    s.emissive = emissive;
    s.n_emissive = 1;

    orc_env e;
    memset(&e, 0, sizeof(e));
    for (int k = 0; k < 3; ++k) { e.top[k] = 1.4f; e.bottom[k] = 0.3f; }
    orc_settings st;
    memset(&st, 0, sizeof(st));
    st.max_bounces = 4;
    st.exposure = 1.0f;
    st.gamma = 2.2f;
    st.selected[0] = st.selected[1] = st.selected[2] = -100;

    static float a1[W * H * 4], a2[W * H * 4];
    static double s1[W * H * 3], s2[W * H * 3];
    double checksum = 0.0;
    for (int lens = 0; lens < 2; ++lens) {
        orc_camera c;
        memset(&c, 0, sizeof(c));
        const float eye[3] = { 900.0f, 700.0f, -1100.0f }, target[3] = { 0, 0, 0 };
        float mv[16];
        orc_camera_matrices(eye, target, 0.8f, 1.0f, 10000.0f, lens, W, H, mv, c.inv_mv, c.proj, c.inv_proj);
        c.viewport[2] = W;
        c.viewport[3] = H;
        c.win_z = 0.55f;
        c.lens_model = lens;
        c.lens_radius = lens ? 20.0f : 0.0f;
        c.focal_distance = 1500.0f;
        orc_counters c1, c2;
        render(&s, &e, &c, &st, threads, a1, s1, &c1);
        render(&s, &e, &c, &st, 1, a2, s2, &c2);
        if (memcmp(a1, a2, sizeof(a1)) || memcmp(s1, s2, sizeof(s1)) || memcmp(&c1, &c2, sizeof(c1))) {
            fprintf(stderr, "tsan_render: %d-thread render differs from the 1-thread render (lens %d)\n", threads,
                    lens);
            return 1;
        }
        render(&s, &e, &c, &st, threads, a2, s2, &c2);
        if (memcmp(a1, a2, sizeof(a1)) || memcmp(&c1, &c2, sizeof(c1))) {
            fprintf(stderr, "tsan_render: two %d-thread renders differ (lens %d)\n", threads, lens);
            return 1;
        }
        for (int i = 0; i < W * H * 3; ++i)
            if (s1[i] == s1[i] && s1[i] - s1[i] == 0.0) checksum += s1[i];   /* finite sums only */
        printf("lens %d: rays %llu visits %llu\n", lens, (unsigned long long)c1.rays, (unsigned long long)c1.visits);
    }
    printf("tsan_render ok threads %d checksum %.6f\n", threads, checksum);
    return 0;
}
