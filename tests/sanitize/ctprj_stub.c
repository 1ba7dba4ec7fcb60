/* Host-only stand-in for the three libctprj entry points the gpuPrj gateway calls
 * (prj_setup, prj_run, ctprj_last_error), so the gateway's own logic -- command
 * dispatch, MATLAB <-> C layout conversion, staging buffers, error paths -- can run
 * under the address and undefined-behaviour sanitizers without a GPU.
 *
 * prj_run copies instead of projecting: FWD writes sino[k] = img[k] + 1 over the first
 * min(n*n, np*prjWidth) entries (0 elsewhere), BWD/FBP writes img[k] = sino[k] - 1
 * likewise; the driver checks the gateway's layout conversions against that. */
#include <stddef.h>
#include <string.h>

#include "ctprj.h"

static ctprj_geom g;
static int configured = 0;
static const char* err = "";

int prj_setup(int n, int prjWidth, int np, int prjFull, float dSize, float effectiveRate,
              float d)
{
    if (n < 1 || prjWidth < 1 || np < 1 || prjFull < 1) {
        err = "stub: bad geometry";
        return CTPRJ_EINVAL;
    }
    g.n = n;
    g.prjWidth = prjWidth;
    g.np = np;
    g.prjFull = prjFull;
    g.dSize = dSize;
    g.effectiveRate = effectiveRate;
    g.d = d;
    configured = 1;
    return 0;
}

int prj_run(float* img, float* sino, int cmd)
{
    const size_t ni = (size_t)g.n * g.n, ns = (size_t)g.np * g.prjWidth;
    const size_t m = ni < ns ? ni : ns;
    size_t k;
    if (!configured) {
        err = "stub: prj_setup has not been called";
        return CTPRJ_ESTATE;
    }
    if (cmd & CTPRJ_FWD_BIT) {
        memset(sino, 0, sizeof(float) * ns);
        for (k = 0; k < m; k++) sino[k] = img[k] + 1.0f;
    } else {
        memset(img, 0, sizeof(float) * ni);
        for (k = 0; k < m; k++) img[k] = sino[k] - 1.0f;
    }
    return 0;
}

const char* ctprj_last_error(void) { return err; }
