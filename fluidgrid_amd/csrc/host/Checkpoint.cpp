// Checkpoint.cpp -- HeterogeneousSimulator::saveState / loadState.
//
// File layout (little-endian, the structs as laid out in fluidgrid_amd.h,
// padding zeroed):
//   "FGSTATE1"                                   8 bytes
//   u32 version (1), u32 numFluids, u32 numBodies, u32 0
//   fg_common_fluid_params, fg_common_rigid_body_params
//   per fluid: fg_fluid_params, u32 numParticles, u32 inflow step counter,
//              u32 has inflow, u32 0, fg_inflow_desc,
//              float4 positions[n], float4 velocities[n] (particle order),
//              float2 (rho, p)[n] (sorted order)
//   per dynamic body (B of each, body order): float4 centre of mass,
//              velocity, orientation, angular momentum
//   u64 checksum of everything before it: FNV-1a over 8-byte little-endian
//              words (offset basis 0xcbf29ce484222325, prime 0x100000001b3),
//              the last word zero-padded
// The arrays are streamed in chunks, so a large state never sits whole in
// host memory.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fluidgrid/HeterogeneousSimulator.h"

namespace Simulation {
namespace Simulators {

namespace {
constexpr char kMagic[8] = {'F', 'G', 'S', 'T', 'A', 'T', 'E', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kChunk = size_t(64) << 20;

struct FluidRecord {
    fg_fluid_params params;
    uint32_t numParticles, emitStep, hasInflow, pad;
    fg_inflow_desc inflow;
};

// Word-wise FNV-1a; bytes arrive in pieces of any length.
class Checksum {
public:
    void add(const void* data, size_t n) {
        const unsigned char* p = static_cast<const unsigned char*>(data);
        while (n) {
            const size_t k = std::min(n, size_t(8) - mFill);
            std::memcpy(mWord + mFill, p, k);
            mFill += k;
            p += k;
            n -= k;
            if (mFill == 8) flush();
        }
    }
    uint64_t value() {
        if (mFill) {
            std::memset(mWord + mFill, 0, 8 - mFill);
            flush();
        }
        return mH;
    }

private:
    void flush() {
        uint64_t w;
        std::memcpy(&w, mWord, 8);
        mH = (mH ^ w) * 0x100000001b3ull;
        mFill = 0;
    }
    uint64_t mH = 0xcbf29ce484222325ull;
    unsigned char mWord[8];
    size_t mFill = 0;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

class Writer {
public:
    explicit Writer(const std::string& path) : mOut(path, std::ios::binary | std::ios::trunc), mPath(path) {
        if (!mOut) throw std::runtime_error("saveState: cannot open " + path);
    }
    void put(const void* p, size_t n) {
        mSum.add(p, n);
        mOut.write(static_cast<const char*>(p), (std::streamsize)n);
        if (!mOut) throw std::runtime_error("saveState: write failed for " + mPath);
    }
    // A device array, through a host bounce buffer of at most kChunk bytes.
    void putDevice(const void* dev, size_t n) {
        std::vector<unsigned char> buf(std::min(n, kChunk));
        for (size_t off = 0; off < n; off += kChunk) {
            const size_t k = std::min(kChunk, n - off);
            hip_check(hipMemcpy(buf.data(), static_cast<const unsigned char*>(dev) + off, k, hipMemcpyDeviceToHost),
                      "saveState");
            put(buf.data(), k);
        }
    }
    void finish() {
        const uint64_t h = mSum.value();
        mOut.write(reinterpret_cast<const char*>(&h), sizeof(h));
        mOut.close();
        if (!mOut) throw std::runtime_error("saveState: write failed for " + mPath);
    }

private:
    std::ofstream mOut;
    std::string mPath;
    Checksum mSum;
};

class Reader {
public:
    explicit Reader(const std::string& path) : mIn(path, std::ios::binary), mPath(path) {
        if (!mIn) throw std::runtime_error("loadState: cannot open " + path);
    }
    void get(void* p, size_t n) {
        mIn.read(static_cast<char*>(p), (std::streamsize)n);
        if ((size_t)mIn.gcount() != n) throw std::runtime_error("loadState: " + mPath + " is truncated");
        mSum.add(p, n);
    }
    // n bytes into a device array (dev null: read and checksum only).
    void getDevice(void* dev, size_t n) {
        std::vector<unsigned char> buf(std::min(n, kChunk));
        for (size_t off = 0; off < n; off += kChunk) {
            const size_t k = std::min(kChunk, n - off);
            get(buf.data(), k);
            if (dev)
                hip_check(hipMemcpy(static_cast<unsigned char*>(dev) + off, buf.data(), k, hipMemcpyHostToDevice),
                          "loadState");
        }
    }
    void finish() {
        const uint64_t want = mSum.value();
        uint64_t h = 0;
        mIn.read(reinterpret_cast<char*>(&h), sizeof(h));
        if (mIn.gcount() != (std::streamsize)sizeof(h)) throw std::runtime_error("loadState: " + mPath + " is truncated");
        if (h != want) throw std::runtime_error("loadState: checksum mismatch in " + mPath);
        if (mIn.peek() != std::char_traits<char>::eof())
            throw std::runtime_error("loadState: trailing bytes in " + mPath);
    }

private:
    std::ifstream mIn;
    std::string mPath;
    Checksum mSum;
};

template <class T>
bool same_bytes(const T& a, const T& b) {
    return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// The rigid-body block has 16 bytes of alignment padding; it is written and
// compared with the padding zeroed.
fg_common_rigid_body_params rigid_params(const void* src) {
    fg_common_rigid_body_params in, o;
    std::memcpy(&in, src, sizeof(in));
    std::memset(&o, 0, sizeof(o));
    o.rb_to_fluid_relation = in.rb_to_fluid_relation;
    o.st_bucket_size = in.st_bucket_size;
    o.dn_bucket_size = in.dn_bucket_size;
    o.st_keyBits = in.st_keyBits;
    o.velocity_limit = in.velocity_limit;
    o.volume_min = in.volume_min;
    o.volume_max = in.volume_max;
    o.global_external_force = in.global_external_force;
    o.simulation_scale = in.simulation_scale;
    o.deltaTimeIntegration = in.deltaTimeIntegration;
    o.particleRadius = in.particleRadius;
    return o;
}
}  // namespace

void HeterogeneousSimulator::saveState(const std::string& path) {
    auto& app = *Application::ApplicationData::instance();
    RigidBody::RigidBodySimulator& rb = *RigidBody::RigidBodySimulator::instance();
    const uint32_t B = (uint32_t)rb.getNumDynamicBodies();
    hip_check(hipDeviceSynchronize(), "saveState sync");
    Writer w(path);
    w.put(kMagic, sizeof(kMagic));
    const uint32_t head[4] = {kVersion, (uint32_t)mFluids.size(), B, 0};
    w.put(head, sizeof(head));
    w.put(&app.getCommonFluidParams(), sizeof(fg_common_fluid_params));
    const fg_common_rigid_body_params rp = rigid_params(&app.getCommonRigidBodyParams());
    w.put(&rp, sizeof(rp));
    for (Fluid::GPUSPHFluid* f : mFluids) {
        FluidRecord r{};
        std::memcpy(&r.params, &f->mFluidParams, sizeof(fg_fluid_params));
        r.numParticles = (uint32_t)f->getNumParticles();
        r.emitStep = f->readEmitStep();
        r.hasInflow = f->hasInflow() ? 1u : 0u;
        if (f->hasInflow()) r.inflow = f->getInflow();
        w.put(&r, sizeof(r));
        const size_t n = r.numParticles;
        w.putDevice(f->getPosition(), sizeof(float4) * n);
        w.putDevice(f->getVelocity(), sizeof(float4) * n);
        w.putDevice(f->getDensityField(), sizeof(float2) * n);
    }
    if (B) {
        std::vector<float4> pos(B), vel(B), orient(B), L(B);
        rb.downloadBodies(pos.data(), vel.data(), orient.data(), L.data());
        for (const auto* a : {&pos, &vel, &orient, &L}) w.put(a->data(), sizeof(float4) * B);
    }
    w.finish();
}

void HeterogeneousSimulator::loadState(const std::string& path) {
    auto& app = *Application::ApplicationData::instance();
    RigidBody::RigidBodySimulator& rb = *RigidBody::RigidBodySimulator::instance();
    const uint32_t B = (uint32_t)rb.getNumDynamicBodies();
    // Pass 1 checks the scene and the checksum without touching the state;
    // pass 2 reads the file again into the device buffers.
    for (int pass = 0; pass < 2; pass++) {
        Reader r(path);
        char magic[8];
        r.get(magic, sizeof(magic));
        if (std::memcmp(magic, kMagic, sizeof(kMagic)) != 0)
            throw std::runtime_error("loadState: " + path + " is not a simulator state file");
        uint32_t head[4];
        r.get(head, sizeof(head));
        if (head[0] != kVersion) throw std::runtime_error("loadState: unsupported state file version");
        if (head[1] != mFluids.size() || head[2] != B)
            throw std::runtime_error("loadState: the file holds " + std::to_string(head[1]) + " fluids and " +
                                     std::to_string(head[2]) + " bodies, the simulator " +
                                     std::to_string(mFluids.size()) + " and " + std::to_string(B));
        fg_common_fluid_params cp;
        fg_common_rigid_body_params rp;
        r.get(&cp, sizeof(cp));
        r.get(&rp, sizeof(rp));
        if (!same_bytes(cp, *reinterpret_cast<const fg_common_fluid_params*>(&app.getCommonFluidParams())))
            throw std::runtime_error("loadState: common fluid parameters differ from the simulator's");
        if (!same_bytes(rp, rigid_params(&app.getCommonRigidBodyParams())))
            throw std::runtime_error("loadState: rigid-body parameters differ from the simulator's");
        std::vector<uint32_t> steps;
        for (size_t i = 0; i < mFluids.size(); i++) {
            Fluid::GPUSPHFluid* f = mFluids[i];
            FluidRecord rec;
            r.get(&rec, sizeof(rec));
            const std::string which = "loadState: fluid " + std::to_string(i) + ": ";
            if (!same_bytes(rec.params, *reinterpret_cast<const fg_fluid_params*>(&f->mFluidParams)))
                throw std::runtime_error(which + "fluid parameters differ");
            if (rec.numParticles != (uint32_t)f->getNumParticles())
                throw std::runtime_error(which + "particle count differs");
            if ((rec.hasInflow != 0) != f->hasInflow() || (rec.hasInflow && !same_bytes(rec.inflow, f->getInflow())))
                throw std::runtime_error(which + "inflow differs");
            const size_t n = rec.numParticles;
            const bool write = pass == 1;
            if (write) hip_check(hipDeviceSynchronize(), "loadState sync");
            r.getDevice(write ? f->getPosition() : nullptr, sizeof(float4) * n);
            r.getDevice(write ? f->getVelocity() : nullptr, sizeof(float4) * n);
            r.getDevice(write ? f->getDensityField() : nullptr, sizeof(float2) * n);
            if (write && f->hasInflow()) f->writeEmitStep(rec.emitStep);
        }
        std::vector<float4> pos(B), vel(B), orient(B), L(B);
        for (auto* a : {&pos, &vel, &orient, &L})
            if (B) r.get(a->data(), sizeof(float4) * B);
        r.finish();
        if (pass == 1 && B) rb.uploadBodies(pos.data(), vel.data(), orient.data(), L.data());
    }
}

}  // namespace Simulators
}  // namespace Simulation
