// websocket.cpp -- see websocket.hpp.
#include "websocket.hpp"

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <sys/socket.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstring>
#include <new>
#include <stdexcept>

namespace ostft_wire {
namespace {

// ---- SHA-1 (FIPS 180-1) and base64, for the handshake only ----
uint32_t rol(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

std::string sha1(const std::string& msg)
{
    uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
    std::string m = msg;
    const uint64_t bits = (uint64_t)msg.size() * 8;
    m.push_back((char)0x80);
    while (m.size() % 64 != 56)
        m.push_back('\0');
    for (int i = 7; i >= 0; --i)
        m.push_back((char)((bits >> (8 * i)) & 0xFF));
    for (size_t off = 0; off < m.size(); off += 64) {
        uint32_t w[80];
        for (int i = 0; i < 16; ++i) {
            const unsigned char* p = reinterpret_cast<const unsigned char*>(m.data() + off + 4 * i);
            w[i] = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
        }
        for (int i = 16; i < 80; ++i)
            w[i] = rol(w[i - 3] ^ w[i - 8] ^ w[i - 14] ^ w[i - 16], 1);
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
        for (int i = 0; i < 80; ++i) {
            uint32_t f, k;
            if (i < 20) {
                f = (b & c) | (~b & d);
                k = 0x5A827999u;
            } else if (i < 40) {
                f = b ^ c ^ d;
                k = 0x6ED9EBA1u;
            } else if (i < 60) {
                f = (b & c) | (b & d) | (c & d);
                k = 0x8F1BBCDCu;
            } else {
                f = b ^ c ^ d;
                k = 0xCA62C1D6u;
            }
            const uint32_t tmp = rol(a, 5) + f + e + k + w[i];
            e = d;
            d = c;
            c = rol(b, 30);
            b = a;
            a = tmp;
        }
        h[0] += a;
        h[1] += b;
        h[2] += c;
        h[3] += d;
        h[4] += e;
    }
    std::string out(20, '\0');
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 4; ++j)
            out[4 * i + j] = (char)((h[i] >> (24 - 8 * j)) & 0xFF);
    return out;
}

std::string base64(const std::string& in)
{
    static const char* tbl = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    size_t i = 0;
    for (; i + 2 < in.size(); i += 3) {
        const uint32_t v = ((uint8_t)in[i] << 16) | ((uint8_t)in[i + 1] << 8) | (uint8_t)in[i + 2];
        out += tbl[(v >> 18) & 63];
        out += tbl[(v >> 12) & 63];
        out += tbl[(v >> 6) & 63];
        out += tbl[v & 63];
    }
    if (i + 1 == in.size()) {
        const uint32_t v = (uint8_t)in[i] << 16;
        out += tbl[(v >> 18) & 63];
        out += tbl[(v >> 12) & 63];
        out += "==";
    } else if (i + 2 == in.size()) {
        const uint32_t v = ((uint8_t)in[i] << 16) | ((uint8_t)in[i + 1] << 8);
        out += tbl[(v >> 18) & 63];
        out += tbl[(v >> 12) & 63];
        out += tbl[(v >> 6) & 63];
        out += '=';
    }
    return out;
}

bool read_exact(int fd, void* buf, size_t n)
{
    char* p = static_cast<char*>(buf);
    while (n > 0) {
        const ssize_t r = ::recv(fd, p, n, 0);
        if (r == 0)
            return false;
        if (r < 0) {
            if (errno == EINTR)
                continue;
            return false;
        }
        p += r;
        n -= (size_t)r;
    }
    return true;
}

bool write_all(int fd, const void* buf, size_t n)
{
    const char* p = static_cast<const char*>(buf);
    while (n > 0) {
        const ssize_t r = ::send(fd, p, n, MSG_NOSIGNAL);
        if (r < 0) {
            if (errno == EINTR)
                continue;
            return false;
        }
        p += r;
        n -= (size_t)r;
    }
    return true;
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

// Read the HTTP upgrade request; return the Sec-WebSocket-Key or "".
std::string handshake_key(int fd)
{
    std::string req;
    char c;
    while (req.size() < 16384) {
        if (!read_exact(fd, &c, 1))
            return "";
        req.push_back(c);
        if (req.size() >= 4 && req.compare(req.size() - 4, 4, "\r\n\r\n") == 0)
            break;
    }
    const std::string lo = lower(req);
    if (lo.compare(0, 4, "get ") != 0 || lo.find("upgrade: websocket") == std::string::npos)
        return "";
    const size_t k = lo.find("sec-websocket-key:");
    if (k == std::string::npos)
        return "";
    size_t b = k + 18;
    while (b < req.size() && (req[b] == ' ' || req[b] == '\t'))
        ++b;
    size_t e = req.find("\r\n", b);
    if (e == std::string::npos)
        return "";
    while (e > b && (req[e - 1] == ' ' || req[e - 1] == '\t'))
        --e;
    return req.substr(b, e - b);
}

}  // namespace

std::string ws_accept_key(const std::string& client_key)
{
    return base64(sha1(client_key + "258EAFA5-E914-47DA-95CA-C5AB0DC85B11"));
}

bool WsConnection::send(int opcode, const void* data, uint64_t len)
{
    unsigned char hdr[10];
    size_t h = 0;
    hdr[h++] = (unsigned char)(0x80 | (opcode & 0x0F));
    if (len < 126) {
        hdr[h++] = (unsigned char)len;
    } else if (len <= 0xFFFF) {
        hdr[h++] = 126;
        hdr[h++] = (unsigned char)(len >> 8);
        hdr[h++] = (unsigned char)len;
    } else {
        hdr[h++] = 127;
        for (int i = 7; i >= 0; --i)
            hdr[h++] = (unsigned char)(len >> (8 * i));
    }
    std::lock_guard<std::mutex> lk(wmu_);
    return write_all(fd_, hdr, h) && write_all(fd_, data, (size_t)len);
}

void WsConnection::shutdown() { ::shutdown(fd_, SHUT_RDWR); }

WsServer::WsServer(int port, std::string host, uint64_t max_message)
    : port_(port), host_(std::move(host)), max_message_(max_message)
{
}

WsServer::~WsServer()
{
    stop();
    for (auto& ct : threads_)
        if (ct.t.joinable())
            ct.t.join();
    if (lfd_ >= 0)
        ::close(lfd_);
}

bool WsServer::listen(std::string& err)
{
    lfd_ = ::socket(AF_INET, SOCK_STREAM, 0);
    if (lfd_ < 0) {
        err = std::strerror(errno);
        return false;
    }
    int one = 1;
    ::setsockopt(lfd_, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
    sockaddr_in a{};
    a.sin_family = AF_INET;
    a.sin_port = htons((uint16_t)port_);
    if (::inet_pton(AF_INET, host_.c_str(), &a.sin_addr) != 1) {
        err = "bad listen address";
        return false;
    }
    if (::bind(lfd_, reinterpret_cast<sockaddr*>(&a), sizeof(a)) != 0 || ::listen(lfd_, 64) != 0) {
        err = std::strerror(errno);
        return false;
    }
    socklen_t al = sizeof(a);
    if (port_ == 0 && ::getsockname(lfd_, reinterpret_cast<sockaddr*>(&a), &al) == 0)
        port_ = ntohs(a.sin_port);  // port 0: the one the system chose
    return true;
}

void WsServer::serve(Handler h)
{
    while (!stopping_) {
        sockaddr_in peer{};
        socklen_t pl = sizeof(peer);
        const int fd = ::accept(lfd_, reinterpret_cast<sockaddr*>(&peer), &pl);
        if (fd < 0) {
            if (stopping_)
                break;
            if (errno == EINTR || errno == ECONNABORTED)
                continue;
            break;
        }
        int one = 1;
        ::setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
        std::lock_guard<std::mutex> lk(cmu_);
        for (auto it = threads_.begin(); it != threads_.end();) {
            if (it->done->load()) {
                it->t.join();
                it = threads_.erase(it);
            } else {
                ++it;
            }
        }
        auto done = std::make_shared<std::atomic<bool>>(false);
        threads_.push_back(ConnThread{std::thread(&WsServer::connection_thread, this, fd, h, done), done});
    }
}

void WsServer::connection_thread(int fd, Handler h, std::shared_ptr<std::atomic<bool>> done)
{
    connection(fd, std::move(h));
    done->store(true);
}

void WsServer::stop()
{
    if (stopping_.exchange(true))
        return;
    if (lfd_ >= 0)
        ::shutdown(lfd_, SHUT_RDWR);
    std::lock_guard<std::mutex> lk(cmu_);
    for (auto& c : conns_)
        c->shutdown();
}

void WsServer::connection(int fd, Handler h)
{
    auto conn = std::make_shared<WsConnection>(fd);
    {
        std::lock_guard<std::mutex> lk(cmu_);
        if (stopping_) {
            ::close(fd);
            return;
        }
        conns_.push_back(conn);
    }
    const std::string key = handshake_key(fd);
    if (key.empty()) {
        const char* bad = "HTTP/1.1 400 Bad Request\r\nContent-Length: 0\r\n\r\n";
        write_all(fd, bad, std::strlen(bad));
    } else {
        const std::string resp = "HTTP/1.1 101 Switching Protocols\r\nUpgrade: websocket\r\n"
                                 "Connection: Upgrade\r\nSec-WebSocket-Accept: "
                                 + ws_accept_key(key) + "\r\n\r\n";
        if (write_all(fd, resp.data(), resp.size())) {
            std::string msg;
            int msg_op = 0;
            for (;;) {
                unsigned char b[2];
                if (!read_exact(fd, b, 2))
                    break;
                const bool fin = b[0] & 0x80;
                const int op = b[0] & 0x0F;
                const bool masked = b[1] & 0x80;
                uint64_t len = b[1] & 0x7F;
                if ((b[0] & 0x70) || !masked) {  // no extensions negotiated; clients must mask
                    const unsigned char close_pe[4] = {0x88, 2, 0x03, 0xEA};  // 1002
                    conn->send(8, close_pe + 2, 2);
                    break;
                }
                if (len == 126) {
                    unsigned char e[2];
                    if (!read_exact(fd, e, 2))
                        break;
                    len = ((uint64_t)e[0] << 8) | e[1];
                } else if (len == 127) {
                    unsigned char e[8];
                    if (!read_exact(fd, e, 8))
                        break;
                    len = 0;
                    for (int i = 0; i < 8; ++i)
                        len = (len << 8) | e[i];
                }
                unsigned char mask[4];
                if (!read_exact(fd, mask, 4))
                    break;
                const bool control = op >= 8;
                // the message so far is at most max_message_ long, so the
                // room left cannot wrap (msg.size() + len could, for a 64-bit len)
                if ((control && (len > 125 || !fin)) || len > max_message_ - msg.size()) {
                    const unsigned char code[2] = {0x03, (unsigned char)(control ? 0xEA : 0xF1)};  // 1002 / 1009
                    conn->send(8, code, 2);
                    break;
                }
                std::string payload;
                try {
                    payload.assign((size_t)len, '\0');
                    if (op == 0)
                        msg.reserve(msg.size() + (size_t)len);
                } catch (const std::bad_alloc&) {
                    const unsigned char code[2] = {0x03, 0xF1};  // 1009: message too big
                    conn->send(8, code, 2);
                    break;
                }
                if (len && !read_exact(fd, &payload[0], (size_t)len))
                    break;
                for (uint64_t i = 0; i < len; ++i)
                    payload[i] ^= (char)mask[i & 3];
                if (op == 8) {  // close: echo the status code, then end
                    conn->send(8, payload.data(), payload.size() >= 2 ? 2 : 0);
                    break;
                }
                if (op == 9) {
                    conn->send(10, payload.data(), payload.size());
                    continue;
                }
                if (op == 10)
                    continue;
                if (op == 1 || op == 2) {
                    if (msg_op != 0) {  // a new data frame inside a fragmented message
                        const unsigned char code[2] = {0x03, 0xEA};
                        conn->send(8, code, 2);
                        break;
                    }
                    msg_op = op;
                    msg = std::move(payload);
                } else if (op == 0) {
                    if (msg_op == 0) {
                        const unsigned char code[2] = {0x03, 0xEA};
                        conn->send(8, code, 2);
                        break;
                    }
                    msg += payload;
                } else {
                    const unsigned char code[2] = {0x03, 0xEA};
                    conn->send(8, code, 2);
                    break;
                }
                if (fin) {
                    const int done_op = msg_op;
                    msg_op = 0;
                    try {
                        h(*conn, done_op, msg);
                    } catch (const std::exception&) {
                        // a handler that throws closes this connection only
                        const unsigned char code[2] = {0x03, 0xF3};  // 1011: internal error
                        conn->send(8, code, 2);
                        break;
                    }
                    msg.clear();
                    if (stopping_)
                        break;
                }
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(cmu_);
        conns_.erase(std::remove(conns_.begin(), conns_.end(), conn), conns_.end());
    }
    ::close(fd);
}

}  // namespace ostft_wire
