// fuzz_wire.cpp -- seeded fuzz of the runner's wire code, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (any report aborts the run):
//
//   * the Cap'n Proto request codec (request_codec.cpp): random requests
//     round-trip bit for bit through encode_request / decode_request;
//     mutated encodings (bit flips, truncations, overwritten pointer words,
//     forged segment tables) and random byte strings decode or fail cleanly;
//     the list-size limit of encode_request is checked before any read;
//   * the WebSocket server (websocket.cpp): a server on a loopback port is
//     sent well-formed, fragmented and malformed upgrade requests and frames
//     (unmasked, reserved bits, oversized 64-bit lengths, control frames too
//     long or fragmented, continuation without a start, junk) from many
//     connections; it must answer or close each one and keep serving.
//
// usage: fuzz_wire SEED CODEC_ITERS WS_CONNECTIONS
#include <arpa/inet.h>
#include <netinet/in.h>
#include <sys/socket.h>
#include <sys/time.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "request_codec.hpp"
#include "websocket.hpp"

using ostft_wire::Request;

namespace {

std::mt19937_64 rng;

uint64_t rnd(uint64_t n) { return n ? rng() % n : 0; }

std::string rand_text(size_t max_len)
{
    std::string s(rnd(max_len + 1), '\0');
    for (auto& c : s)
        c = (char)(32 + rnd(95));
    return s;
}

[[noreturn]] void fail(const char* what, uint64_t iter)
{
    std::fprintf(stderr, "FAIL: %s (iteration %llu)\n", what, (unsigned long long)iter);
    std::exit(1);
}

Request rand_request(std::vector<float>& data)
{
    Request r;
    r.sharedMemory = rand_text(rnd(4) ? 40 : 300);
    r.mappedID = rand_text(24);
    r.options = rand_text(rnd(4) ? 80 : 600);
    r.memPTR = rng();
    r.posixFileDes = (int64_t)rng();
    r.windowsHandlePTR = rng();
    r.windowRadix = (uint32_t)rng();
    const uint32_t bits = (uint32_t)rng();
    std::memcpy(&r.overlapRatio, &bits, 4);  // any bit pattern, NaNs included
    r.dataLength = rng();
    r.overlapdataLength = rng();
    data.resize(rnd(4) ? rnd(64) : rnd(5000));
    for (auto& v : data) {
        const uint32_t b = (uint32_t)rng();
        std::memcpy(&v, &b, 4);
    }
    return r;
}

bool same_bits(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

void codec_round_trip(uint64_t it)
{
    std::vector<float> data;
    const Request r = rand_request(data);
    std::string enc;
    if (!ostft_wire::encode_request(r, data.data(), data.size(), enc))
        fail("encode_request refused a valid request", it);
    Request d;
    std::string err;
    if (!ostft_wire::decode_request(reinterpret_cast<const uint8_t*>(enc.data()), enc.size(), d, err))
        fail(("decode of an encoded request failed: " + err).c_str(), it);
    if (d.sharedMemory != r.sharedMemory || d.mappedID != r.mappedID || d.options != r.options || d.memPTR != r.memPTR
        || d.posixFileDes != r.posixFileDes || d.windowsHandlePTR != r.windowsHandlePTR
        || d.windowRadix != r.windowRadix || !same_bits(d.overlapRatio, r.overlapRatio)
        || d.dataLength != r.dataLength || d.overlapdataLength != r.overlapdataLength || d.data.size() != data.size())
        fail("round trip changed a field", it);
    for (size_t i = 0; i < data.size(); ++i)
        if (!same_bits(d.data[i], data[i]))
            fail("round trip changed a data value", it);
}

void decode_any(const std::string& m)
{
    Request d;
    std::string err;
    (void)ostft_wire::decode_request(reinterpret_cast<const uint8_t*>(m.data()), m.size(), d, err);
}

void codec_mutations(uint64_t it)
{
    std::vector<float> data;
    const Request r = rand_request(data);
    std::string enc;
    if (!ostft_wire::encode_request(r, data.data(), data.size(), enc))
        fail("encode_request refused a valid request", it);
    std::string m = enc;
    const int kind = (int)rnd(6);
    if (kind == 0) {  // bit flips
        for (uint64_t k = 0, n = 1 + rnd(8); k < n && !m.empty(); ++k)
            m[rnd(m.size())] ^= (char)(1u << rnd(8));
    } else if (kind == 1) {  // truncation
        m.resize(rnd(m.size()));
    } else if (kind == 2) {  // overwrite a word (pointers, sizes, the segment table)
        if (m.size() >= 8) {
            const size_t w = rnd(m.size() / 8) * 8;
            const uint64_t v = rnd(2) ? rng() : (rng() & 0xFFFFFFFFull) | ((uint64_t)rnd(4) << 0);
            std::memcpy(&m[w], &v, 8);
        }
    } else if (kind == 3) {  // forged segment table: many segments / huge sizes
        if (m.size() >= 8) {
            const uint32_t segs = (uint32_t)(rnd(2) ? rng() : rnd(16));
            std::memcpy(&m[0], &segs, 4);
            const uint32_t sz = (uint32_t)(rnd(2) ? rng() : rnd(1u << 20));
            std::memcpy(&m[4], &sz, 4);
        }
    } else if (kind == 4) {  // appended junk
        for (uint64_t k = 0, n = rnd(64); k < n; ++k)
            m.push_back((char)rng());
    } else {  // a double-far or far pointer at a random word
        if (m.size() >= 16) {
            const size_t w = (1 + rnd(m.size() / 8 - 1)) * 8;
            const uint64_t v = 2u | (rnd(2) << 2) | (rnd(1u << 20) << 3) | ((uint64_t)rnd(8) << 32);
            std::memcpy(&m[w], &v, 8);
        }
    }
    decode_any(m);
}

void codec_random_bytes()
{
    std::string m(rnd(4) ? rnd(64) : rnd(4096), '\0');
    for (auto& c : m)
        c = (char)rng();
    decode_any(m);
}

void codec_limits()
{
    Request r;
    std::string out;
    const float one = 1.0f;
    // a list of 2^29 elements does not fit: refused before the data is read
    if (ostft_wire::encode_request(r, &one, 1ull << 29, out))
        fail("encode_request accepted a 2^29-element list", 0);
    if (ostft_wire::encode_request(r, &one, ~0ull, out))
        fail("encode_request accepted a 2^64-1-element list", 0);
    decode_any(std::string());
    decode_any(std::string(7, '\0'));
}

// ---- WebSocket ----

int connect_to(int port)
{
    const int fd = ::socket(AF_INET, SOCK_STREAM, 0);
    if (fd < 0)
        return -1;
    timeval tv{0, 200000};
    ::setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &tv, sizeof(tv));
    ::setsockopt(fd, SOL_SOCKET, SO_SNDTIMEO, &tv, sizeof(tv));
    sockaddr_in a{};
    a.sin_family = AF_INET;
    a.sin_port = htons((uint16_t)port);
    ::inet_pton(AF_INET, "127.0.0.1", &a.sin_addr);
    if (::connect(fd, reinterpret_cast<sockaddr*>(&a), sizeof(a)) != 0) {
        ::close(fd);
        return -1;
    }
    return fd;
}

void send_all(int fd, const std::string& s)
{
    size_t off = 0;
    while (off < s.size()) {
        const ssize_t r = ::send(fd, s.data() + off, s.size() - off, MSG_NOSIGNAL);
        if (r <= 0)
            return;  // the server may close early: that is an allowed answer
        off += (size_t)r;
    }
}

void drain(int fd)
{
    char buf[4096];
    for (int k = 0; k < 64; ++k)
        if (::recv(fd, buf, sizeof(buf), 0) <= 0)
            return;
}

std::string frame(bool fin, int rsv, int op, bool masked, uint64_t len_field, int len_form, const std::string& payload)
{
    std::string f;
    f.push_back((char)((fin ? 0x80 : 0) | ((rsv & 7) << 4) | (op & 0x0F)));
    const char mbit = masked ? (char)0x80 : 0;
    if (len_form == 0) {
        f.push_back((char)(mbit | (char)(len_field & 0x7F)));
    } else if (len_form == 1) {
        f.push_back((char)(mbit | 126));
        f.push_back((char)(len_field >> 8));
        f.push_back((char)len_field);
    } else {
        f.push_back((char)(mbit | 127));
        for (int i = 7; i >= 0; --i)
            f.push_back((char)(len_field >> (8 * i)));
    }
    unsigned char mask[4] = {(unsigned char)rng(), (unsigned char)rng(), (unsigned char)rng(), (unsigned char)rng()};
    if (masked)
        f.append(reinterpret_cast<const char*>(mask), 4);
    for (size_t i = 0; i < payload.size(); ++i)
        f.push_back(masked ? (char)(payload[i] ^ (char)mask[i & 3]) : payload[i]);
    return f;
}

std::string good_frame(bool fin, int op, const std::string& p)
{
    const uint64_t n = p.size();
    return frame(fin, 0, op, true, n, n < 126 ? 0 : n <= 0xFFFF ? 1 : 2, p);
}

std::string handshake(bool valid)
{
    if (!valid) {
        static const char* bad[] = {"POST / HTTP/1.1\r\n\r\n", "GET / HTTP/1.1\r\nHost: x\r\n\r\n",
                                    "GET / HTTP/1.1\r\nUpgrade: websocket\r\n\r\n",
                                    "GET / HTTP/1.1\r\nUpgrade: websocket\r\nSec-WebSocket-Key:\r\n\r\n"};
        if (rnd(3))
            return bad[rnd(4)];
        std::string junk(rnd(20000), '\0');  // over the 16 KiB header limit, or short junk
        for (auto& c : junk)
            c = (char)(rnd(2) ? 'a' + rnd(26) : rng());
        return junk;
    }
    return "GET /x HTTP/1.1\r\nHost: 127.0.0.1\r\nUpgrade: websocket\r\nConnection: Upgrade\r\n"
           "Sec-WebSocket-Key: dGhlIHNhbXBsZSBub25jZQ==\r\nSec-WebSocket-Version: 13\r\n\r\n";
}

std::string rand_frames()
{
    std::string s;
    for (uint64_t k = 0, n = 1 + rnd(6); k < n; ++k) {
        const int kind = (int)rnd(12);
        const std::string p = rand_text(rnd(8) ? 40 : 70000);
        switch (kind) {
        case 0: s += good_frame(true, 1, p); break;                        // text
        case 1: s += good_frame(true, 2, p); break;                        // binary
        case 2:                                                            // fragmented
            s += good_frame(false, 2, p) + good_frame(false, 0, p) + good_frame(true, 0, p);
            break;
        case 3: s += good_frame(true, 9, p.substr(0, 100)); break;          // ping (maybe > 125)
        case 4: s += good_frame(false, 9, "x"); break;                     // fragmented control
        case 5: s += good_frame(true, 0, p); break;                        // continuation, no start
        case 6: s += frame(true, 0, 2, false, p.size() < 126 ? p.size() : 0, 0, p); break;  // unmasked
        case 7: s += frame(true, 1 + (int)rnd(7), 1, true, 3, 0, "abc"); break;  // reserved bits
        case 8: s += frame(true, 0, 2, true, (1ull << 63) + rnd(1000), 2, ""); break;  // huge length
        case 9: s += frame(true, 0, 2, true, 2u << 20, 2, ""); break;      // over max_message
        case 10: s += good_frame(true, 3 + (int)rnd(5), p); break;         // reserved opcode
        default: {                                                         // junk
            std::string j(rnd(300), '\0');
            for (auto& c : j)
                c = (char)rng();
            s += j;
        }
        }
    }
    if (rnd(2))
        s += good_frame(true, 8, std::string("\x03\xe8", 2));  // close
    return s;
}

void ws_fuzz(uint64_t conns)
{
    ostft_wire::WsServer srv(0, "127.0.0.1", 1u << 20);
    std::string err;
    if (!srv.listen(err))
        fail(("listen: " + err).c_str(), 0);
    std::atomic<uint64_t> messages{0};
    std::thread th([&] {
        srv.serve([&](ostft_wire::WsConnection& c, int op, std::string& payload) {
            messages++;
            if (payload.size() >= 5 && payload.compare(0, 5, "throw") == 0)
                throw std::runtime_error("handler error");
            c.send(op, payload.data(), payload.size() < 64 ? payload.size() : 64);
        });
    });
    for (uint64_t i = 0; i < conns; ++i) {
        const int fd = connect_to(srv.port());
        if (fd < 0)
            fail("connect", i);
        const bool valid = rnd(5) != 0;
        send_all(fd, handshake(valid));
        if (valid) {
            if (rnd(10) == 0)
                send_all(fd, good_frame(true, 1, "throw now"));
            send_all(fd, rand_frames());
        }
        if (rnd(2))
            ::shutdown(fd, SHUT_WR);
        drain(fd);
        ::close(fd);
    }
    // the server still answers a clean exchange
    const int fd = connect_to(srv.port());
    if (fd < 0)
        fail("connect after fuzz", conns);
    send_all(fd, handshake(true));
    send_all(fd, good_frame(true, 1, "hello"));
    std::string got;
    char buf[512];
    for (int k = 0; k < 50 && got.find("hello") == std::string::npos; ++k) {
        const ssize_t r = ::recv(fd, buf, sizeof(buf), 0);
        if (r > 0)
            got.append(buf, (size_t)r);
        else if (r == 0)
            break;
    }
    ::close(fd);
    if (got.find("101 Switching Protocols") == std::string::npos || got.find("hello") == std::string::npos)
        fail("server does not answer after the fuzz", conns);
    srv.stop();
    th.join();
    std::printf("websocket: %llu connections, %llu messages handled\n", (unsigned long long)conns,
                (unsigned long long)messages.load());
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s SEED CODEC_ITERS WS_CONNECTIONS\n", argv[0]);
        return 2;
    }
    rng.seed(std::strtoull(argv[1], nullptr, 10));
    const uint64_t iters = std::strtoull(argv[2], nullptr, 10), conns = std::strtoull(argv[3], nullptr, 10);
    codec_limits();
    for (uint64_t i = 0; i < iters; ++i) {
        codec_round_trip(i);
        codec_mutations(i);
        codec_random_bytes();
    }
    std::printf("codec: %llu round trips, %llu mutated and %llu random messages decoded\n",
                (unsigned long long)iters, (unsigned long long)iters, (unsigned long long)iters);
    ws_fuzz(conns);
    std::printf("ok\n");
    return 0;
}
