// request_codec.hpp -- Cap'n Proto wire codec for the reference's request
// struct, written against the published encoding (flat-array framing,
// struct/list/far pointers), without the Cap'n Proto library.
//
// Schema (StandAlone/include/FFTcapnp.capnp:1-16), struct RequestCapnp with a
// 6-word data section and 4 pointers (StandAlone/include/FFTcapnp.capnp.h:35):
//
//   data bytes  0  memPTR            UInt64      pointer 0  sharedMemory  Text
//               8  posixFileDes      Int64       pointer 1  data          List(Float32)
//              16  windowsHandlePTR  UInt64      pointer 2  mappedID      Text
//              24  windowRadix       UInt32      pointer 3  options       Text
//              28  overlapRatio      Float32
//              32  dataLength        UInt64
//              40  overlapdataLength UInt64
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ostft_wire {

struct Request {
    std::string sharedMemory;
    std::string mappedID;
    std::string options;
    uint64_t memPTR = 0;
    int64_t posixFileDes = 0;
    uint64_t windowsHandlePTR = 0;
    uint32_t windowRadix = 0;
    float overlapRatio = 0.0f;
    uint64_t dataLength = 0;
    uint64_t overlapdataLength = 0;
    std::vector<float> data;  // the List(Float32), empty when absent
};

// Decode one flat-array message (segment table + segments).  Accepts any
// number of segments, single- and double-far pointers, and struct sections
// smaller or larger than the schema's.  Never reads outside `buf`; on a
// malformed message returns false with a reason in `err`.
bool decode_request(const uint8_t* buf, size_t len, Request& out, std::string& err);

// Encode `r` as a single-segment flat array; `data`/`n` are written as the
// data list (n may be 0: no list).  Returns false if n does not fit a
// Cap'n Proto list (2^29 - 1 elements).
bool encode_request(const Request& r, const float* data, uint64_t n, std::string& out);

}  // namespace ostft_wire
