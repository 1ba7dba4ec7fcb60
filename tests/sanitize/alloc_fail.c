/* malloc with an armed failure, for the gateway's host staging allocations: the
 * gateway is compiled with -Dmalloc=mock_malloc and this file without it. */
#include <stdlib.h>

static int g_countdown = 0;

/* the k-th mock_malloc call from now returns NULL (k > 0); 0 disarms */
void mock_fail_malloc(int k) { g_countdown = k; }

void* mock_malloc(size_t n)
{
    if (g_countdown > 0 && --g_countdown == 0) return NULL;
    return malloc(n);
}
