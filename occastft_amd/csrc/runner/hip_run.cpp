// hip_run.cpp -- `hipRun`, the MI355X runner process.
//
// A drop-in for the reference runner executables (serialRun, openmpRun,
// cudaRun): `hipRun [port]` listens on 0.0.0.0:port (default 54500) for
// WebSocket connections and answers Cap'n Proto RequestCapnp messages, with
// the input either in the message's data list or in the POSIX shared memory
// object the request names (RunnerCommon.cpp:61-164, FFTStruct_linux.cpp).
// The reply copies the request's fields (without options), carries the result
// in the data list or writes it into the shared memory, and marks a failed
// request with sharedMemory = "ERR" (FFTStruct.cpp:15-22).  A text message
// "CLOSE_REQUEST" is answered with "CLOSE_COMPLETE" and stops the server.
//
// Differences from the reference runners, on purpose: every request keeps its
// own shared-memory mapping (the reference shares one `dataInfo` member across
// connection threads, RunnerInterface.hpp:120); an undecodable message is
// answered with an "ERR" reply instead of an exception; a result larger than
// the client's mapping is an error instead of an overflow.
#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/occastft_hip.h"
#include "request_codec.hpp"
#include "websocket.hpp"

namespace {

using ostft_wire::Request;

// A Cap'n Proto list holds fewer than 2^29 elements: the largest result the
// data list can carry.
constexpr uint64_t MAX_LIST_FLOATS = (1ull << 29) - 1;

// page-rounded byte size of n floats (FFTStruct.hpp:138-145); 0 when it does
// not fit in 64 bits
uint64_t page_bytes(uint64_t nfloats)
{
    const uint64_t page = (uint64_t)sysconf(_SC_PAGESIZE);
    if (nfloats > (UINT64_MAX - page) / sizeof(float))
        return 0;
    const uint64_t bytes = nfloats * sizeof(float);
    return ((bytes + page - 1) / page) * page;
}

struct ShmMap {
    void* ptr = nullptr;
    int fd = -1;
    uint64_t bytes = 0;
    ~ShmMap()
    {
        if (ptr && ptr != MAP_FAILED)
            munmap(ptr, bytes);
        if (fd >= 0)
            close(fd);
    }
};

// The runner: the reference's Runner shape (RunnerInterface.hpp:108-131) with
// the HIP engine behind ActivateSTFT.
class HipRunner {
public:
    explicit HipRunner(int port) : server_(port) {}
    ~HipRunner()
    {
        if (init_thread_.joinable())
            init_thread_.join();
        UnInit();
    }

    // Listens first and brings the device up on a thread of its own: the
    // reference proxy gives a runner five one-second connection attempts
    // (STFTProxy.cpp:136-145), less than a cold device's start-up, so the
    // WebSocket handshake must not wait for it.  Requests that arrive earlier
    // wait in ActivateSTFT until the environment is ready.
    bool ServerInit(std::string& err)
    {
        if (!server_.listen(err))
            return false;
        init_thread_ = std::thread([this] {
            InitEnv();
            BuildKernel();
            std::lock_guard<std::mutex> lk(ready_mu_);
            ready_ = true;
            ready_cv_.notify_all();
        });
        return true;
    }
    void Serve()
    {
        server_.serve([this](ostft_wire::WsConnection& c, int op, std::string& msg) { OnMessage(c, op, msg); });
    }

    // Runner::ActivateSTFT: returns false where the reference returns nullopt.
    // Requests go to the continuous-batching server, so requests of
    // concurrent connections share batches on the GPU; a request whose input
    // and output do not fit in the server's arena runs on its own through a
    // context of its own.
    // `max_out`: the most floats the reply can carry (the client's shared
    // memory, or a data list); a larger result is refused before anything is
    // allocated for it.
    bool ActivateSTFT(const std::vector<float>& in, int exp, float overlap, const std::string& options,
                      uint64_t max_out, std::vector<float>& out, std::string& why)
    {
        // the request itself is checked first: it needs no device
        ostft_params p;
        int rc = ostft_parse_options(options.c_str(), exp, overlap, &p);
        ostft_geometry g;
        if (rc == OSTFT_OK)
            rc = ostft_get_geometry(&p, in.size(), &g);
        if (rc != OSTFT_OK) {
            why = ostft_strerror(rc);
            return false;
        }
        if (g.out_floats > max_out) {
            why = "result of " + std::to_string(g.out_floats) + " floats exceeds what the reply carries ("
                  + std::to_string(max_out) + ")";
            return false;
        }
        {
            std::unique_lock<std::mutex> lk(ready_mu_);
            ready_cv_.wait(lk, [this] { return ready_; });
        }
        if (!srv_) {
            why = "no HIP device";
            return false;
        }
        out.resize(g.out_floats ? g.out_floats : 1);
        uint64_t ticket = 0, n = 0;
        rc = ostft_server_submit(srv_, in.empty() ? nullptr : in.data(), in.size(), &p, out.data(), out.size(),
                                 &ticket);
        if (rc == OSTFT_OK) {
            rc = ostft_server_wait(srv_, ticket, -1, &n);
            if (rc != OSTFT_OK)
                why = std::string(ostft_strerror(rc)) + ": " + ostft_server_last_error(srv_);
        } else if (rc == OSTFT_ERR_TOO_LARGE) {
            std::lock_guard<std::mutex> lk(big_mu_);
            if (!big_ctx_ && ostft_create(0, &big_ctx_) != OSTFT_OK)
                big_ctx_ = nullptr;
            rc = big_ctx_ ? ostft_activate(big_ctx_, in.empty() ? nullptr : in.data(), in.size(), exp, overlap,
                                           options.c_str(), out.data(), out.size(), &n)
                          : OSTFT_ERR_NO_DEVICE;
            if (rc != OSTFT_OK)
                why = std::string(ostft_strerror(rc)) + (big_ctx_ ? std::string(": ") + ostft_last_error(big_ctx_)
                                                                  : std::string());
        } else {
            why = std::string(ostft_strerror(rc)) + ": " + ostft_server_last_error(srv_);
        }
        if (rc != OSTFT_OK)
            return false;
        out.resize(n);
        return true;
    }

private:
    void InitEnv()
    {
        ostft_server_config cfg;
        std::memset(&cfg, 0, sizeof(cfg));
        cfg.use_graphs = 1;
        if (const char* a = std::getenv("OSTFT_ARENA_MB"))
            cfg.arena_bytes = (uint64_t)std::strtoull(a, nullptr, 10) << 20;
        if (ostft_server_create(0, &cfg, &srv_) != OSTFT_OK) {
            srv_ = nullptr;
            std::cerr << "hipRun: no HIP device; every request will be answered with ERR" << std::endl;
        }
    }
    void BuildKernel() {}  // kernels are compiled ahead of time into the engine library
    void UnInit()
    {
        if (srv_) {
            ostft_server_stats st;
            if (ostft_server_stats_get(srv_, &st, 0) == OSTFT_OK && st.completed > 0)
                std::fprintf(stderr,
                             "hipRun: %llu requests in %llu batches, latency p50 %.3f ms p99 %.3f ms\n",
                             (unsigned long long)st.completed, (unsigned long long)st.batches, st.latency_p50_ms,
                             st.latency_p99_ms);
            ostft_server_destroy(srv_);
        }
        srv_ = nullptr;
        if (big_ctx_)
            ostft_destroy(big_ctx_);
        big_ctx_ = nullptr;
    }

    // RunnerCommon.cpp AccessData: the input from shm or from the data list.
    bool AccessData(const Request& req, std::vector<float>& data, ShmMap& shm, std::string& why)
    {
        if (!req.sharedMemory.empty()) {
            shm.fd = shm_open(req.sharedMemory.c_str(), O_RDWR, 0666);
            if (shm.fd < 0) {
                why = "shm_open failed";
                return false;
            }
            shm.bytes = page_bytes(req.overlapdataLength);
            if (shm.bytes == 0 || req.dataLength > shm.bytes / sizeof(float)) {
                why = "shared memory smaller than the input";
                return false;
            }
            shm.ptr = mmap(nullptr, shm.bytes, PROT_READ | PROT_WRITE, MAP_SHARED, shm.fd, 0);
            if (shm.ptr == MAP_FAILED) {
                shm.ptr = nullptr;
                why = "mmap failed";
                return false;
            }
            data.resize(req.dataLength);
            std::memcpy(data.data(), shm.ptr, req.dataLength * sizeof(float));
            return true;
        }
        if (req.data.empty()) {
            why = "no input data";
            return false;
        }
        data = req.data;
        return true;
    }

    void OnMessage(ostft_wire::WsConnection& conn, int opcode, std::string& msg)
    {
        if (opcode == 1) {
            if (msg == "CLOSE_REQUEST") {
                conn.send_text("CLOSE_COMPLETE");
                std::thread([this] { server_.stop(); }).detach();
            }
            return;
        }
        Request rep;
        std::string out;
        try {
            out = Handle(msg, rep);
        } catch (const std::exception& e) {
            // an allocation or a library call that throws ends this request
            // with an ERR reply, not the runner (other clients keep their
            // connections)
            std::cerr << "hipRun: request " << rep.mappedID << " failed: " << e.what() << std::endl;
            rep.sharedMemory = "ERR";
            out.clear();
            ostft_wire::encode_request(rep, nullptr, 0, out);
        }
        conn.send_binary(out);
    }

    // Decodes one binary message and runs it; returns the encoded reply.
    // `rep` receives the reply fields as soon as they are known.
    std::string Handle(std::string& msg, Request& rep)
    {
        Request req;
        std::string err;
        const bool decoded
            = ostft_wire::decode_request(reinterpret_cast<const uint8_t*>(msg.data()), msg.size(), req, err);
        msg.clear();
        msg.shrink_to_fit();
        std::vector<float> result;
        bool ok = decoded;
        if (decoded) {
            // FFTRequest::MakeWField: every field but options and data
            rep.windowRadix = req.windowRadix;
            rep.overlapRatio = req.overlapRatio;
            rep.mappedID = req.mappedID;
            rep.sharedMemory = req.sharedMemory;
            rep.memPTR = req.memPTR;
            rep.posixFileDes = req.posixFileDes;
            rep.windowsHandlePTR = req.windowsHandlePTR;
            rep.dataLength = req.dataLength;
            rep.overlapdataLength = req.overlapdataLength;
            ShmMap shm;
            std::vector<float> data;
            ok = AccessData(req, data, shm, err);
            req.data.clear();
            req.data.shrink_to_fit();
            const uint64_t max_out = shm.ptr ? shm.bytes / sizeof(float) : MAX_LIST_FLOATS;
            if (ok)
                ok = ActivateSTFT(data, (int)req.windowRadix, req.overlapRatio, req.options, max_out, result, err);
            if (ok && shm.ptr) {
                if (result.size() > shm.bytes / sizeof(float)) {
                    ok = false;
                    err = "result larger than the shared memory";
                } else {
                    std::memcpy(shm.ptr, result.data(), result.size() * sizeof(float));
                    result.clear();
                }
            }
        }
        if (!ok) {
            std::cerr << "hipRun: request " << rep.mappedID << " failed: " << err << std::endl;
            rep.sharedMemory = "ERR";
            result.clear();
        } else if (!result.empty()) {
            rep.dataLength = result.size();  // FFTRequest::SetData
        }
        std::string out;
        if (!ostft_wire::encode_request(rep, result.data(), result.size(), out)) {
            rep.sharedMemory = "ERR";
            out.clear();
            ostft_wire::encode_request(rep, nullptr, 0, out);
        }
        return out;
    }

    ostft_wire::WsServer server_;
    std::thread init_thread_;
    std::mutex ready_mu_;
    std::condition_variable ready_cv_;
    bool ready_ = false;  // InitEnv has finished (srv_ set, or null: no device)
    ostft_server* srv_ = nullptr;
    ostft_ctx* big_ctx_ = nullptr;  // requests larger than the server's arena
    std::mutex big_mu_;
};

}  // namespace

// `hipRun --decode FILE [OUT]`: print a request message's fields as JSON
// (and, with OUT, write it back re-encoded as one segment) -- a diagnostic for
// client authors.
int decode_file(const char* path, const char* out_path)
{
    FILE* fh = std::fopen(path, "rb");
    if (!fh) {
        std::perror(path);
        return 1;
    }
    std::string buf;
    char chunk[1 << 16];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof(chunk), fh)) > 0)
        buf.append(chunk, r);
    std::fclose(fh);
    Request q;
    std::string err;
    if (!ostft_wire::decode_request(reinterpret_cast<const uint8_t*>(buf.data()), buf.size(), q, err)) {
        std::printf("{\"ok\": false, \"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    double sum = 0.0;
    for (float v : q.data)
        sum += v;
    auto js = [](const std::string& s) {
        std::string o;
        for (char c : s) {
            if (c == '"' || c == '\\')
                o += '\\';
            if ((unsigned char)c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof(b), "\\u%04x", (unsigned char)c);
                o += b;
                continue;
            }
            o += c;
        }
        return o;
    };
    std::printf("{\"ok\": true, \"sharedMemory\": \"%s\", \"mappedID\": \"%s\", \"options\": \"%s\", "
                "\"memPTR\": %llu, \"posixFileDes\": %lld, \"windowsHandlePTR\": %llu, \"windowRadix\": %u, "
                "\"overlapRatio\": %.9g, \"dataLength\": %llu, \"overlapdataLength\": %llu, "
                "\"data_count\": %zu, \"data_sum\": %.17g}\n",
                js(q.sharedMemory).c_str(), js(q.mappedID).c_str(), js(q.options).c_str(),
                (unsigned long long)q.memPTR, (long long)q.posixFileDes, (unsigned long long)q.windowsHandlePTR,
                q.windowRadix, (double)q.overlapRatio, (unsigned long long)q.dataLength,
                (unsigned long long)q.overlapdataLength, q.data.size(), sum);
    if (out_path) {
        std::string out;
        if (!ostft_wire::encode_request(q, q.data.data(), q.data.size(), out))
            return 1;
        FILE* oh = std::fopen(out_path, "wb");
        if (!oh || std::fwrite(out.data(), 1, out.size(), oh) != out.size())
            return 1;
        std::fclose(oh);
    }
    return 0;
}

int main(int argc, char* argv[])
{
    if (argc >= 3 && std::strcmp(argv[1], "--decode") == 0)
        return decode_file(argv[2], argc >= 4 ? argv[3] : nullptr);
    int port = 54500;  // RunnerCommon.cpp:146
    if (argc == 2) {
        try {
            port = std::stoi(argv[1]);
        } catch (const std::exception& e) {
            std::cerr << e.what() << '\n';
            return 1;
        }
    }
    HipRunner runner(port);
    std::string err;
    if (!runner.ServerInit(err)) {
        std::cerr << "ERR on server init: " << err << std::endl;
        return 1;
    }
    std::printf("hipRun listening on %d\n", port);
    std::fflush(stdout);
    runner.Serve();
    return 0;
}
