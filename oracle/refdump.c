/*
 * ORACLE — test infrastructure only.
 *
 * LD_PRELOAD interposer for the reference program (oracle/_ref/mersenne2_ref),
 * written against the public OpenCL API only; the reference itself is run
 * unmodified. The reference prints no residue: after its last iteration it
 * reads the digit state back with one blocking clEnqueueReadBuffer
 * (mersenne2_gpu.cpp:4855-4862: h elements of {u64 s0, s1; u32 t0, t1}, or
 * 2h u64 digits in its compact PRP cycle, :4857-4860) and only tests it. This
 * library passes every call through and, for blocking reads, also writes the
 * bytes read to the file named by M2_REF_DUMP (each blocking read replaces
 * the file, so the last one, the final state, remains). The caller turns the
 * digits into the canonical residue (oracle/run_reference.py).
 *
 * With M2_REF_CPU=1 it also maps the reference's GPU device enumeration
 * (clGetDeviceIDs with CL_DEVICE_TYPE_GPU, :3743-3752) onto CPU devices, for
 * a CPU OpenCL platform such as PoCL where one is installed.
 */
#define _GNU_SOURCE
#define CL_TARGET_OPENCL_VERSION 120
#include <CL/cl.h>
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>

typedef cl_int (*read_fn)(cl_command_queue, cl_mem, cl_bool, size_t, size_t, void *, cl_uint, const cl_event *,
                          cl_event *);
typedef cl_int (*devs_fn)(cl_platform_id, cl_device_type, cl_uint, cl_device_id *, cl_uint *);

CL_API_ENTRY cl_int CL_API_CALL clEnqueueReadBuffer(cl_command_queue q, cl_mem buf, cl_bool blocking, size_t offset,
                                                    size_t size, void *ptr, cl_uint nwait, const cl_event *wait,
                                                    cl_event *ev) {
    static read_fn real = NULL;
    if (!real) real = (read_fn)dlsym(RTLD_NEXT, "clEnqueueReadBuffer");
    if (!real) return CL_INVALID_OPERATION;
    cl_int rc = real(q, buf, blocking, offset, size, ptr, nwait, wait, ev);
    const char *path = getenv("M2_REF_DUMP");
    if (rc == CL_SUCCESS && blocking && path && *path && size > 0) {
        char tmp[4096];
        snprintf(tmp, sizeof tmp, "%s.tmp", path);
        FILE *f = fopen(tmp, "wb");
        if (f) {
            size_t w = fwrite(ptr, 1, size, f);
            if (fclose(f) == 0 && w == size) rename(tmp, path);
        }
    }
    return rc;
}

CL_API_ENTRY cl_int CL_API_CALL clGetDeviceIDs(cl_platform_id platform, cl_device_type type, cl_uint num,
                                               cl_device_id *devices, cl_uint *num_devices) {
    static devs_fn real = NULL;
    if (!real) real = (devs_fn)dlsym(RTLD_NEXT, "clGetDeviceIDs");
    if (!real) return CL_INVALID_OPERATION;
    const char *cpu = getenv("M2_REF_CPU");
    if (cpu && cpu[0] == '1' && type == CL_DEVICE_TYPE_GPU) type = CL_DEVICE_TYPE_CPU;
    return real(platform, type, num, devices, num_devices);
}
