// Host check of the product's boundary-shell maps (seis::Shell): the save
// enumeration and the inject lookup are inverse maps over the same cells and
// every cell lies inside the padded grid.  Built and run by test_lib_cpu.py.
#include <cstdio>
#include <vector>

#include "seis_internal.hpp"

int main() {
    int bad = 0, cases = 0;
    for (int nd = 2; nd <= 3; ++nd)
        for (int H = 1; H <= 6; ++H)
            for (int fs = 0; fs <= 1; ++fs)
                for (int nab : {4, 9}) {
                    const int NZ = 37, NY = nd == 3 ? 33 : 1, NX = 41;
                    seis::Layout g = seis::make_layout(nd, NZ, NY, NX, H, 4);
                    seis::Shell sh = seis::make_shell(g, nab, fs);
                    const int NZp = NZ + 2 * H, NYp = nd == 3 ? NY + 2 * H : 1, NXp = NX + 2 * H;
                    std::vector<char> seen(sh.n, 0);
                    for (long gid = 0; gid < sh.n; ++gid) {
                        int k, j, i;
                        sh.cell(gid, k, j, i);
                        if (k < H || k >= NZp - H || i < H || i >= NXp - H ||
                            (nd == 3 && (j < H || j >= NYp - H)))
                            ++bad;
                        else if (sh.inject(k, j, i) != gid)
                            ++bad;
                    }
                    long inside = 0;
                    for (int i = 0; i < NXp; ++i)
                        for (int j = 0; j < NYp; ++j)
                            for (int k = 0; k < NZp; ++k) {
                                const long m = sh.inject(k, j, i);
                                if (m < 0) continue;
                                ++inside;
                                if (m >= sh.n || seen[m]) ++bad;
                                else seen[m] = 1;
                            }
                    if (inside != sh.n) ++bad;
                    ++cases;
                }
    std::printf("cases %d bad %d\n", cases, bad);
    return bad != 0;
}
