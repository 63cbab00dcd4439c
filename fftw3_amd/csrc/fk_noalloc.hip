/*
 * Static device scratch for the allocation-free execute
 * (fftw_amd_execute_noalloc in api.c, fk_execute_noalloc in engine.c).
 *
 * `scratch` is a file-scope device array of FK_STATIC_SCRATCH_BYTES; each
 * device that loads this code object holds its own copy until the process
 * ends.  Host code never uses the array's name as an address (on the host
 * it names a stand-in, not device memory): fk_static_scratch_setup takes
 * the device address with hipGetSymbolAddress once per device, with that
 * device current, and every later use goes through that address.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "engine.h"

#define FK_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[FK_STATIC_SCRATCH_BYTES];

static void *g_scratch_ptr[FK_NOALLOC_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

/* Takes the current device's address of `scratch` if not yet taken.
   Returns 0 or a HIP error code. */
extern "C" int fk_static_scratch_setup(void)
{
    int dev = -1, rc = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= FK_NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    pthread_mutex_lock(&g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = 0;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p)
            g_scratch_ptr[dev] = p;
        else
            rc = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
    }
    pthread_mutex_unlock(&g_scratch_mu);
    return rc;
}

/* Device address of the current device's `scratch` (set up on first use
   if the library's init ran on another device).  Returns 0 or a HIP error
   code; *bytes = FK_STATIC_SCRATCH_BYTES. */
extern "C" int fk_static_scratch(void **ptr, size_t *bytes)
{
    int dev = -1, rc;
    hipError_t e = hipGetDevice(&dev);
    *ptr = 0;
    *bytes = FK_STATIC_SCRATCH_BYTES;
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= FK_NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    pthread_mutex_lock(&g_scratch_mu);
    *ptr = g_scratch_ptr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    if (*ptr) return 0;
    rc = fk_static_scratch_setup();
    if (rc) return rc;
    pthread_mutex_lock(&g_scratch_mu);
    *ptr = g_scratch_ptr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    return 0;
}
