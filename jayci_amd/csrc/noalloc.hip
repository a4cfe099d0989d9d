// Allocation-free sigma entry point: jci_sigma_dev with every device buffer beyond the
// caller's x and y taken from one file-scope device array instead of hipMalloc.
//
// The array is JCI_NOALLOC_SCRATCH_BYTES (128 MiB) per GPU, held from the library's load on
// that GPU until the process ends.  Its device address is looked up once per device when a
// context is created on it (jci_noalloc_setup, called by jci_ctx_create with the device
// current) and only that address is used afterwards.
//
// What the array holds per call: the FP32 copy of x (FP32 methods), and for the dense-string
// sigma the task descriptors, the opposite-spin intermediate E (La * N elements per vector)
// and the vector staging of nvec > 1.  Every buffer is written before it is read within a
// call, so nothing needs clearing between calls.  All calls on one GPU share the array:
// calls that overlap on one GPU (two host threads, or streams running concurrently) race.
#include "jci_internal.h"

#include <mutex>

#define JCI_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
#define JCI_NOALLOC_MAXDEV 64

static __device__ __align__(256) unsigned char scratch[JCI_NOALLOC_SCRATCH_BYTES];

static std::mutex g_mu;
static void *g_ptr[JCI_NOALLOC_MAXDEV];

// once per device, with that device current
int jci_noalloc_setup(int device)
{
    if (device < 0 || device >= JCI_NOALLOC_MAXDEV) {
        jci_set_error("noalloc: device %d out of range", device);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        jci_set_error("noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s",
                      device, hipGetErrorString(e));
        return -2;
    }
    g_ptr[device] = p;
    return 0;
}

JCI_API long long jci_sigma_noalloc_capacity(void) { return (long long)JCI_NOALLOC_SCRATCH_BYTES; }

JCI_API int jci_sigma_dev_noalloc(jci_ham *h, int method, int nvec, const double *x_dev,
                                  double *y_dev)
{
    if (!h || !x_dev || !y_dev) { jci_set_error("sigma_dev_noalloc: null argument"); return -1; }
    const int dev = jci_ham_device(h);
    void *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev >= 0 && dev < JCI_NOALLOC_MAXDEV) p = g_ptr[dev];
    }
    if (!p) { jci_set_error("sigma_dev_noalloc: scratch array not set up on device %d", dev); return -2; }
    SigmaScratch s{(unsigned char *)p, JCI_NOALLOC_SCRATCH_BYTES, 0};
    return jci_sigma_dev_scratch(h, method, nvec, x_dev, y_dev, &s);
}
