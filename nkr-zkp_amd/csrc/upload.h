// Host-to-device witness upload for proofs whose witness is host memory
// (Backend::prove takes a std::vector<Fr>, protocols/groth16/common/include/
// groth16/backend.hpp:23-27; the reference's HIP backend copies it whole,
// backends/hip/src/backend.hip:302-307).
//
// Witnesses are mostly zeros and ones (bits, selectors, sparse inputs), and a
// 2^20-column witness is 32 MiB, which takes longer to copy through one host
// thread into pinned memory and over PCIe than the GPU takes to prove a
// sparse circuit.  So the witness travels packed:
//   * host worker threads scan 1 MiB chunks in parallel; per group of 64
//     elements they record a mask of the elements that are neither 0 nor 1,
//     a mask of those equal to 1 (Montgomery form R mod r) and where the
//     group's other values start, and pack those values into pinned memory;
//   * the calling thread issues each chunk's host-to-device copy as soon as
//     the chunk is packed, so packing and PCIe overlap;
//   * one kernel expands the groups into the device witness (one wave per
//     group: a lane finds its value by a popcount of the mask below it).
// A dense witness moves as many bytes as a plain copy, still pipelined and
// packed on several threads.  The device witness is bit-identical to the host
// one.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

#include "field.h"
#include "runtime.h"
#include "witness_pack.h"

namespace nkr {

class WitnessUpload {
public:
    explicit WitnessUpload(uint32_t cols);
    ~WitnessUpload();
    WitnessUpload(const WitnessUpload &) = delete;
    WitnessUpload &operator=(const WitnessUpload &) = delete;

    // Enqueues on `stream` the transfer of host[0 .. cols) into z_dev (device
    // memory of cols elements); returns once every copy is issued.  The host
    // array may be reused when this returns.  Uploads through one object must
    // be serialised (the caller's context is not thread-safe either).
    void upload(const fr_t *host, fr_t *z_dev, hipStream_t stream);

    // Bytes sent over PCIe by the last upload (packed values, masks, offsets).
    uint64_t last_bytes() const { return last_bytes_; }
    uint64_t device_bytes() const { return pack_dev_.bytes() + meta_dev_.bytes(); }

    // Elements per chunk handed to one host worker (32768 = 1 MiB).
    static constexpr uint32_t kChunk = 1u << 15;
    static constexpr uint32_t kGroup = kWitnessGroup;

private:
    uint32_t cols_, groups_, chunks_;
    // Pinned: packed values (cols), then the group metadata -- masks (2 per
    // group: values other than 0 and 1, values equal to 1) followed by the
    // offsets (1 per group) -- which the device holds in the same layout.
    fr_t *vals_host_ = nullptr;
    uint8_t *meta_host_ = nullptr;
    std::atomic<uint32_t> *chunk_count_ = nullptr;  // per chunk: packed values + 1 once packed (0 = pending)
    DeviceBuffer<fr_t> pack_dev_;
    DeviceBuffer<uint8_t> meta_dev_;
    size_t meta_bytes_ = 0;
    hipEvent_t done_ = nullptr;  // the last upload's copies have been read
    uint64_t last_bytes_ = 0;
};

// Several witnesses in one packed transfer (the batch prover's upload).  The
// witnesses are laid end to end in one index space, each starting at a
// multiple of 64 elements, and packed by the same rule as one witness; every
// chunk of at most kChunk elements belongs to one witness.  One kernel
// expands all chunks, each into its own witness's device array.  Buffers grow
// to the largest batch and are reused.
class WitnessBatchUpload {
public:
    struct Item {
        const fr_t *host;
        fr_t *z_dev;
        uint32_t cols;
    };
    WitnessBatchUpload() = default;
    WitnessBatchUpload(const WitnessBatchUpload &) = delete;
    WitnessBatchUpload &operator=(const WitnessBatchUpload &) = delete;

    // Enqueues on `stream` every item's transfer and its expansion; returns
    // once every copy is issued.  The host arrays may be reused then.
    void upload(const std::vector<Item> &items, hipStream_t stream);
    uint64_t last_bytes() const { return last_bytes_; }
    void release_device();

    // Where each chunk's expansion writes (the last table of the metadata).
    struct Chunk {
        fr_t *z;         // the witness's device array, at the chunk's first element
        uint32_t begin;  // first element in the shared index space (a multiple of 64)
        uint32_t count;  // elements, at most kChunk
    };

private:
    PinnedStaging<fr_t> vals_;
    PinnedStaging<uint8_t> meta_;
    DeviceBuffer<fr_t> pack_dev_;
    DeviceBuffer<uint8_t> meta_dev_;
    std::unique_ptr<std::atomic<uint32_t>[]> chunk_count_;
    size_t chunk_cap_ = 0;
    uint64_t last_bytes_ = 0;
};

}  // namespace nkr
