#include "device_app.hpp"

namespace gat {

DeviceApp::DeviceApp(int device) {
    owned_.reset(new Context());
    if (owned_->init(device) == ST_OK) {
        ctx_ = owned_.get();
    } else {
        owned_.reset();
    }
}

DeviceApp::DeviceApp(Context* shared) : ctx_(shared) {}

DeviceApp::~DeviceApp() { releaseBuffers(); }

int DeviceApp::alloc(size_t bytes, const char* name) {
    if (!ctx_) return -1;
    Buf b;
    b.bytes = bytes;
    b.name = name ? name : "";
    b.host.assign(bytes ? bytes : 1, 0);
    if (ctx_->alloc(bytes, &b.dev) != ST_OK) return -1;
    bufs_.push_back(std::move(b));
    return (int)bufs_.size() - 1;
}

int DeviceApp::createRaw(size_t bytes, const char* name) { return alloc(bytes, name); }

bool DeviceApp::syncToDevice(int h) {
    ctx_->make_current();
    Buf& b = bufs_[h];
    if (hipMemcpyAsync(b.dev, b.host.data(), b.bytes, hipMemcpyHostToDevice, ctx_->stream) != hipSuccess)
        return false;
    return hipStreamSynchronize(ctx_->stream) == hipSuccess;
}

bool DeviceApp::syncFromDevice(int h) {
    ctx_->make_current();
    Buf& b = bufs_[h];
    if (hipMemcpyAsync(b.host.data(), b.dev, b.bytes, hipMemcpyDeviceToHost, ctx_->stream) != hipSuccess)
        return false;
    return hipStreamSynchronize(ctx_->stream) == hipSuccess;
}

bool DeviceApp::clearBuffer(int h) {
    Buf& b = bufs_[h];
    std::memset(b.host.data(), 0, b.bytes);
    return syncToDevice(h);
}

void DeviceApp::releaseBuffers() {
    if (!ctx_) return;
    for (Buf& b : bufs_)
        if (b.dev) ctx_->release(b.dev);
    bufs_.clear();
}

bool DeviceApp::wait() {
    if (!ctx_) return false;
    ctx_->make_current();
    return hipStreamSynchronize(ctx_->stream) == hipSuccess;
}

}  // namespace gat
