// Host-only probe of hostpot::build_fe_table for the CPU tests: reads
//   nv av[nv] rv[nv] np ap[np] rp[np]
// from stdin, runs set_up and the table build, and prints
//   ok binh binc, then the MDK_FE_TAB table words (records as float and as
//   int bit patterns), then B: the breakpoints the table was built on.
#include <stdio.h>
#include <string.h>

#include "../../moldy_amd/csrc/mdk_hostpot.h"

int main() {
  hostpot::Knots k;
  memset(&k, 0, sizeof(k));
  if (scanf("%d", &k.nv) != 1) return 2;
  for (int i = 0; i < k.nv; i++) scanf("%lf", &k.av[i]);
  for (int i = 0; i < k.nv; i++) scanf("%lf", &k.rv[i]);
  if (scanf("%d", &k.np) != 1) return 2;
  for (int i = 0; i < k.np; i++) scanf("%lf", &k.ap[i]);
  for (int i = 0; i < k.np; i++) scanf("%lf", &k.rp[i]);
  hostpot::set_up(k);
  static float tab[MDK_FE_TAB];
  float binh = 0.f, binc = 0.f;
  const bool ok = hostpot::build_fe_table(k, tab, binh, binc);
  printf("%d %.9g %.9g %.17g\n", ok ? 1 : 0, binh, binc, k.pcr);
  for (int t = 0; t < MDK_FE_TAB; t++) {
    int bits;
    memcpy(&bits, &tab[t], sizeof(int));
    printf("%.9g %d\n", tab[t], bits);
  }
  return 0;
}
