/* Memory and undefined-behaviour check of the oracle (test infrastructure):
 * built with -fsanitize=address,undefined by `make -C oracle asan`, run by
 * tests/test_oracle_asan.py on a system file it writes (int32 n_water,
 * n_surf, n_carbons; then float64 positions (N, 3), velocities v dt (N, 3),
 * carbons (n_carbons, 3)).  Drives every entry point of sesmd_oracle.h in the
 * reference geometry and in a 16^3-box domain at p = 8, so that any
 * out-of-bounds access or undefined operation on those paths aborts the run. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sesmd_oracle.h"
#include "ff_reference.h"

_Static_assert(sizeof(orc_ff) == sizeof(sesmd_forcefield), "orc_ff mirrors sesmd_forcefield");

static int nw, ns, ngr;
static double *r0, *v0, *gr;

static int run(int nb2, int p)
{
    const int n = 3 * nw + 29 * ns;
    sesmd_forcefield sf;
    memset(&sf, 0, sizeof sf);
    sesmd_reference_forcefield(&sf);
    orc_ff ff;
    memcpy(&ff, &sf, sizeof ff);
    double *r = malloc(3 * n * sizeof(double)), *v = malloc(3 * n * sizeof(double));
    double *a = calloc(3 * n, sizeof(double)), *f = calloc(3 * n, sizeof(double));
    double *fs = calloc(3 * n, sizeof(double));
    memcpy(r, r0, 3 * n * sizeof(double));
    memcpy(v, v0, 3 * n * sizeof(double));
    void *h = orc_create(nw, ns, nb2, p, &ff);
    if (!h) return 1;
    orc_set_state(h, r, v, a, 0.0);
    orc_set_nh(h, 3.81087, 1e-4, -2.0 * 100 * 3.81087);
    orc_set_graphite(h, ngr, gr);
    orc_rebuild(h);
    orc_fmm(h);
    orc_forces(h, ORC_ALL, ~0ull, f);
    orc_forces(h, ORC_NEAR, 0x5555555555555ull, f);
    orc_init_accel(h, f);
    double e[ORC_E_COUNT], sums[6];
    orc_energies(h, e);
    orc_move_a(h, sums);
    orc_rebuild(h);
    orc_fmm(h);
    orc_forces(h, ORC_ALL, ~0ull, f);
    orc_move_b(h, f);
    int kb = 8;
    orc_steps(h, 12, 10, &kb, fs);
    orc_set_xi(h, orc_xi(h));
    int nb = 4 * nb2, nmol = nw + ns, lp = (p + 1) * (p + 1);
    int *cls = calloc(nmol, sizeof(int)), *anc = calloc(nmol, sizeof(int));
    int *cbox = calloc(n, sizeof(int)), *rk = calloc(n, sizeof(int));
    unsigned *isub = calloc(n, sizeof(unsigned));
    orc_get_classes(h, cls, anc, cbox, rk, isub);
    size_t n4 = (size_t)nb * nb * nb * lp * 2, n3 = n4 / 8, n2 = (size_t)nb2 * nb2 * nb2 * lp * 2;
    double *x4 = malloc(n4 * 8), *y4 = malloc(n4 * 8), *x3 = malloc(n3 * 8), *y3 = malloc(n3 * 8);
    double *x2 = malloc(n2 * 8), *y2 = malloc(n2 * 8);
    orc_get_expansions(h, x4, y4, x3, y3, x2, y2);
    double ylm[2];
    orc_ylm(h, p, 1, 0.7, 0.3, ylm);
    double eta;
    orc_get_state(h, r, v, a, &eta);
    printf("nb2 %d p %d: kb %d eta %.6e shake failures %d, U %.6e\n", nb2, p, kb, eta, orc_shake_failures(h),
           e[ORC_E_POTENTIAL]);
    orc_destroy(h);
    free(r), free(v), free(a), free(f), free(fs), free(cls), free(anc), free(cbox), free(rk), free(isub);
    free(x4), free(y4), free(x3), free(y3), free(x2), free(y2);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: asan_driver system.bin\n");
        return 2;
    }
    FILE *fh = fopen(argv[1], "rb");
    int hdr[3];
    if (!fh || fread(hdr, sizeof(int), 3, fh) != 3) return 2;
    nw = hdr[0], ns = hdr[1], ngr = hdr[2];
    size_t n = 3 * (size_t)nw + 29 * (size_t)ns;
    r0 = malloc(3 * n * sizeof(double)), v0 = malloc(3 * n * sizeof(double));
    gr = malloc((3 * (size_t)ngr + 1) * sizeof(double));
    if (fread(r0, sizeof(double), 3 * n, fh) != 3 * n || fread(v0, sizeof(double), 3 * n, fh) != 3 * n ||
        fread(gr, sizeof(double), 3 * (size_t)ngr, fh) != 3 * (size_t)ngr)
        return 2;
    fclose(fh);
    orc_set_threads(1);
    if (run(3, 4) || run(4, 8)) return 1;
    free(r0), free(v0), free(gr);
    printf("asan driver ok\n");
    return 0;
}
