// libcusimp.so: drop-in for the library the reference's Blender add-on loads
// (pamo_blender/__init__.py:75-79, built from pamo_blender/cusimpinterface.cu).
// It exports pamo_simplify with the add-on's host-pointer signature and
// forwards to pamo_blender_simplify of libpamo_hip.so, loaded from this
// library's own directory with RTLD_LOCAL so that libpamo_hip.so's own
// pamo_simplify (a different signature) stays separate.
//
// As in the reference, one process-wide stage-2 context (device 0) serves
// every call, so its persistent buffers carry across calls
// (cusimpinterface.cu:55, `static cusimp_free::CUSimp_Free pamo`).  Errors are
// printed to stderr as the reference prints CUDA errors; the counts in
// in_out_nv_ni are then set to 0 so that no stale output is read.
#include <dlfcn.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "pamo_blender.h"

extern "C" {
typedef struct pamo_ctx pamo_ctx;
}

namespace {

struct Api {
    void* lib = nullptr;
    int (*create)(pamo_ctx**, int) = nullptr;
    int (*run)(pamo_ctx*, const float*, const int*, float*, int*, int*, float) = nullptr;
    const char* (*last_error)() = nullptr;
    pamo_ctx* ctx = nullptr;
    std::string err;
};

Api& api() {
    static Api a;
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info;
        std::string dir = ".";
        if (dladdr(reinterpret_cast<void*>(&api), &info) && info.dli_fname) {
            dir = info.dli_fname;
            const size_t k = dir.find_last_of('/');
            dir = k == std::string::npos ? "." : dir.substr(0, k);
        }
        const std::string path = dir + "/libpamo_hip.so";
        a.lib = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!a.lib) {
            a.err = std::string("cannot load ") + path + ": " + dlerror();
            return;
        }
        a.create = reinterpret_cast<int (*)(pamo_ctx**, int)>(dlsym(a.lib, "pamo_ctx_create"));
        a.run = reinterpret_cast<int (*)(pamo_ctx*, const float*, const int*, float*, int*, int*, float)>(
            dlsym(a.lib, "pamo_blender_simplify"));
        a.last_error = reinterpret_cast<const char* (*)()>(dlsym(a.lib, "pamo_last_error"));
        if (!a.create || !a.run || !a.last_error) {
            a.err = path + " lacks pamo_ctx_create / pamo_blender_simplify / pamo_last_error";
            a.run = nullptr;
            return;
        }
        if (a.create(&a.ctx, 0) != 0) {
            a.err = std::string("pamo_ctx_create: ") + a.last_error();
            a.run = nullptr;
        }
    });
    return a;
}

}  // namespace

extern "C" __attribute__((visibility("default"))) void pamo_simplify(float* in_verts, int* in_tris, float* out_verts,
                                                                     int* out_tris, int* in_out_nv_ni,
                                                                     float threshold) {
    Api& a = api();
    if (!a.run) {
        std::fprintf(stderr, "pamo_simplify: %s\n", a.err.c_str());
        if (in_out_nv_ni) in_out_nv_ni[0] = in_out_nv_ni[1] = 0;
        return;
    }
    if (a.run(a.ctx, in_verts, in_tris, out_verts, out_tris, in_out_nv_ni, threshold) != 0) {
        std::fprintf(stderr, "pamo_simplify: %s\n", a.last_error());
        if (in_out_nv_ni) in_out_nv_ni[0] = in_out_nv_ni[1] = 0;
    }
}
