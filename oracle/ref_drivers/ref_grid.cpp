/* Driver for the reference grid builder: prints every array of
 * Gridn::mk_grid(im, jm, km, im0, sm) (src/montecarlo/definitions.h:68-136)
 * as raw float bits, one array per line.  The reference's compile-time
 * Lx, Ly, Lz and fact_mesh (src/param/param.h:13-16) apply.
 * usage: ref_grid im jm km im0 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include "param.h"
#include "definitions.h"

static void dump(const char *name, const float *a, int n) {
    std::printf("%s %d", name, n);
    for (int i = 0; i < n; ++i) { uint32_t b; std::memcpy(&b, &a[i], 4); std::printf(" %08x", b); }
    std::printf("\n");
}

int main(int argc, char **argv) {
    if (argc != 5) { std::fprintf(stderr, "usage: ref_grid im jm km im0\n"); return 2; }
    const int im = std::atoi(argv[1]), jm = std::atoi(argv[2]), km = std::atoi(argv[3]), im0 = std::atoi(argv[4]);
    Gridn g;
    g.mk_grid(im, jm, km, im0, 1);
    dump("x", g.x, im + 2);
    dump("xu", g.xu, im + 1);
    dump("y", g.y, jm + 2);
    dump("yu", g.yu, jm + 1);
    dump("z", g.z, km + 2);
    dump("zu", g.zu, km + 1);
    std::printf("Lx %a Ly %a Lz %a fact_mesh %a\n", (double)Lx, (double)Ly, (double)Lz, (double)fact_mesh);
    g.destroyVar();
    return 0;
}
