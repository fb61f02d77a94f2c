// websocket.hpp -- a small RFC 6455 WebSocket server (the transport of the
// reference runners, which use IXWebSocket v11.4.5; see
// StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:61-142).
//
// One thread per connection.  Client frames must be masked; fragmented
// messages are reassembled; pings are answered; extensions (permessage-deflate,
// offered by the reference client) are declined, which RFC 6455 allows.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace ostft_wire {

class WsConnection {
public:
    explicit WsConnection(int fd) : fd_(fd) {}
    bool send_text(const std::string& s) { return send(1, s.data(), s.size()); }
    bool send_binary(const std::string& s) { return send(2, s.data(), s.size()); }
    bool send(int opcode, const void* data, uint64_t len);
    void shutdown();
    int fd() const { return fd_; }

private:
    int fd_;
    std::mutex wmu_;
};

class WsServer {
public:
    // opcode 1 (text) or 2 (binary), with the whole reassembled payload
    using Handler = std::function<void(WsConnection&, int opcode, std::string& payload)>;

    WsServer(int port, std::string host = "0.0.0.0", uint64_t max_message = (16ull << 30));
    ~WsServer();
    bool listen(std::string& err);
    void serve(Handler h);  // blocks until stop()
    void stop();            // safe from any thread, including a handler
    int port() const { return port_; }  // after listen(): the bound port (also for port 0)

private:
    void connection(int fd, Handler h);
    int port_;
    std::string host_;
    uint64_t max_message_;
    int lfd_ = -1;
    std::atomic<bool> stopping_{false};
    std::mutex cmu_;
    std::vector<std::shared_ptr<WsConnection>> conns_;
    // connection threads; a finished one (done set as its last step) is
    // joined at the next accept, so a long-running server keeps only live ones
    struct ConnThread {
        std::thread t;
        std::shared_ptr<std::atomic<bool>> done;
    };
    std::vector<ConnThread> threads_;
    void connection_thread(int fd, Handler h, std::shared_ptr<std::atomic<bool>> done);
};

// exposed for tests
std::string ws_accept_key(const std::string& client_key);

}  // namespace ostft_wire
