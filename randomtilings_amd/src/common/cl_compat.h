// cl_compat.h — the small slice of the OpenCL C++ API that RandomTilings callers touch
// before constructing a tiler (e.g. Examples/MinimalDominoEx.cpp:17-25):
//
//     cl::Context context(CL_DEVICE_TYPE_DEFAULT);
//     std::vector<cl::Device> devices;
//     context.getInfo(CL_CONTEXT_DEVICES, &devices);
//     devices[0].getInfo(CL_DEVICE_NAME, &sinfo);
//     cl::CommandQueue queue(context);
//     cl_int err = 0;
//
// Here a Context names HIP devices, a Device is one HIP device and a CommandQueue is the
// device a tiler will run on (each tiler owns its own HIP stream).  There is no OpenCL
// dependency and nothing is compiled at run time.
#ifndef RT_CL_COMPAT_H_
#define RT_CL_COMPAT_H_

#include <stdint.h>

#include <string>
#include <vector>

typedef int32_t cl_int;
typedef uint64_t cl_device_type;
typedef uint32_t cl_context_info;
typedef uint32_t cl_device_info;

#ifndef CL_SUCCESS
#define CL_SUCCESS 0
#endif
#define CL_DEVICE_TYPE_DEFAULT (1u << 0)
#define CL_DEVICE_TYPE_CPU (1u << 1)
#define CL_DEVICE_TYPE_GPU (1u << 2)
#define CL_DEVICE_TYPE_ALL 0xFFFFFFFFu
#define CL_CONTEXT_DEVICES 0x1081u
#define CL_DEVICE_NAME 0x102Bu

namespace cl {

class Device {
   public:
    Device() : id_(0) {}
    explicit Device(int id) : id_(id) {}
    int id() const { return id_; }
    // CL_DEVICE_NAME: the HIP device name and its gfx target
    cl_int getInfo(cl_device_info what, std::string* out) const;

   private:
    int id_;
};

class Context {
   public:
    explicit Context(cl_device_type type = CL_DEVICE_TYPE_DEFAULT);
    // CL_CONTEXT_DEVICES: every visible HIP device
    cl_int getInfo(cl_context_info what, std::vector<Device>* out) const;
    int device() const { return device_; }

   private:
    int device_;
};

class CommandQueue {
   public:
    CommandQueue() : device_(0) {}
    explicit CommandQueue(const Context& ctx) : device_(ctx.device()) {}
    CommandQueue(const Context& ctx, const Device& dev) : device_(dev.id()) { (void)ctx; }
    int device() const { return device_; }

   private:
    int device_;
};

}  // namespace cl

#endif  // RT_CL_COMPAT_H_
