// Host build of the hoisted key switch's workgroup placement (hoist_slot in
// cipher-pu_amd/csrc/launch.h): for every tile size and both XCD settings,
// over grids of every shape the library launches (and ragged ones), each
// workgroup index maps into [0, l+1) x [0, R) x [0, batch) and no slot is
// hit twice.  Exit status 0 when every placement is a bijection.
#include <cstdio>
#include <vector>

#include "launch.h"

int main() {
    long shapes = 0, failures = 0;
    for (int l = 1; l <= 6; ++l)
        for (int R = 1; R <= 32; R += (R < 4 ? 1 : 5))
            for (int batch : {1, 2, 3, 7, 8, 19, 64, 65}) {
                const int G = (l + 1) * R * batch;
                for (int xcd = 0; xcd < 2; ++xcd)
                    for (int btile = 1; btile <= batch; btile += (btile < 8 ? 1 : 7)) {
                        ++shapes;
                        std::vector<char> hit(G, 0);
                        bool ok = true;
                        for (int i = 0; i < G && ok; ++i) {
                            const TRB s = hoist_slot(i, G, R, batch, HoistOrder{btile, xcd});
                            if (s.t < 0 || s.t > l || s.r < 0 || s.r >= R || s.b < 0 || s.b >= batch) {
                                ok = false;
                                break;
                            }
                            char& h = hit[((size_t)s.t * R + s.r) * batch + s.b];
                            if (h) ok = false;
                            h = 1;
                        }
                        if (!ok && failures++ < 8)
                            std::printf("FAIL l=%d R=%d batch=%d btile=%d xcd=%d\n", l, R, batch, btile, xcd);
                    }
            }
    std::printf("%ld placements, %ld failures\n", shapes, failures);
    return failures ? 1 : 0;
}
