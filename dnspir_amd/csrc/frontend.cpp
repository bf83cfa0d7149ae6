// DNS-over-PIR front end: the reference server's TCP protocol in front of the
// serving loop.
//
// Frames and commands follow crates/dns-over-pir/src/dns/protocol.rs: every
// message is a 4-byte big-endian length and that many payload bytes (at most
// 64 MiB, :105-118); a request is [cmd][body] with CMD_INFO = 1, CMD_QUERY = 2,
// CMD_DUMP = 3 (:44-53); an unavailable capability answers the single byte
// 0xFF (:85). A connection is handled as the reference's `handle`
// (server.rs:747-807) does it: frames are read in a loop until the peer
// closes, INFO answers flags / num_entries / zone name / dump size / salt
// (encode_info, protocol.rs:166-175), QUERY answers the reply bytes of the
// query (here through dnspir_server_submit / _wait instead of process_query,
// so queries of all connections are batched together), DUMP answers the
// gzipped zone; an empty request or an unknown command ends the connection
// with an error. The listener runs one thread per connection, as `run`
// (server.rs:152-173) does.
#include <arpa/inet.h>
#include <errno.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <string.h>
#include <sys/socket.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <list>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

using namespace dne;

namespace {

constexpr uint8_t CMD_INFO = 1, CMD_QUERY = 2, CMD_DUMP = 3;  // protocol.rs:44-53
constexpr uint8_t FLAG_PIR = 1, FLAG_DUMP = 2;                // protocol.rs:55-60
constexpr uint8_t RESP_UNAVAILABLE = 0xFF;                    // protocol.rs:85
constexpr size_t FRAME_MAX = (size_t)64 << 20;                // protocol.rs:111

// reads exactly n bytes: 1 = read, 0 = the stream ended first, < 0 = error
int read_full(int fd, uint8_t* p, size_t n) {
    size_t got = 0;
    while (got < n) {
        const ssize_t r = ::read(fd, p + got, n - got);
        if (r > 0) got += (size_t)r;
        else if (r == 0) return 0;
        else if (errno != EINTR) return DNSPIR_E_IO;
    }
    return 1;
}

int write_full(int fd, const uint8_t* p, size_t n) {
    size_t put = 0;
    bool sock = true;
    while (put < n) {
        // MSG_NOSIGNAL: a peer that went away is an error, not a SIGPIPE
        ssize_t r = sock ? ::send(fd, p + put, n - put, MSG_NOSIGNAL) : ::write(fd, p + put, n - put);
        if (r < 0 && errno == ENOTSOCK && sock) {
            sock = false;
            continue;
        }
        if (r > 0) put += (size_t)r;
        else if (r == 0 || errno != EINTR) return DNSPIR_E_IO;
    }
    return 0;
}

struct Conn {
    std::thread th;
    int fd = -1;
    std::atomic<bool> done{false};
};

}  // namespace

struct dnspir_frontend {
    dnspir_server* srv = nullptr;
    int db = 0;
    uint32_t num_entries = 0;
    std::string zone;
    uint64_t salt = 0;
    bool has_dump = false;
    std::vector<uint8_t> dump;
    std::atomic<uint64_t> next_client{1};
    std::atomic<uint64_t> connections{0}, frames_in{0}, frames_out{0}, bytes_in{0}, bytes_out{0}, queries{0}, infos{0},
        dumps{0}, unavailable{0}, errors{0};
    // listener
    std::mutex mu;
    int lfd = -1;
    bool stopping = false;
    std::thread acceptor;
    std::list<std::unique_ptr<Conn>> conns;

    int write_frame(int fd, const uint8_t* p, size_t n) {
        const uint8_t len[4] = {(uint8_t)(n >> 24), (uint8_t)(n >> 16), (uint8_t)(n >> 8), (uint8_t)n};
        int rc = write_full(fd, len, 4);
        if (!rc && n) rc = write_full(fd, p, n);
        if (!rc) {
            ++frames_out;
            bytes_out += 4 + n;
        }
        return rc;
    }

    // encode_info (protocol.rs:166-175)
    std::vector<uint8_t> info() const {
        std::vector<uint8_t> o;
        auto be32 = [&](uint32_t v) {
            for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
        };
        o.push_back((uint8_t)((srv ? FLAG_PIR : 0) | (has_dump ? FLAG_DUMP : 0)));
        be32(srv ? num_entries : 0);
        be32((uint32_t)zone.size());
        o.insert(o.end(), zone.begin(), zone.end());
        be32(has_dump ? (uint32_t)dump.size() : 0);
        for (int s = 56; s >= 0; s -= 8) o.push_back((uint8_t)(salt >> s));
        return o;
    }

    // the reference's `handle` (server.rs:747-807) on one connection
    int handle(int fd) {
        ++connections;
        const uint64_t client = next_client++;  // the connection's key-cache identity
        std::vector<uint8_t> req, reply(REPLY_BYTES);
        for (;;) {
            uint8_t lb[4];
            int r = read_full(fd, lb, 4);
            if (r == 0) return 0;  // the peer closed between commands
            if (r < 0) return r;
            const size_t len = ((size_t)lb[0] << 24) | ((size_t)lb[1] << 16) | ((size_t)lb[2] << 8) | lb[3];
            if (len > FRAME_MAX) return DNSPIR_E_FRAME;  // "frame too large"
            req.resize(len);
            r = len ? read_full(fd, req.data(), len) : 1;
            // a frame cut short ends the connection as the peer closing does
            // (read_exact's UnexpectedEof, which `handle` treats as a clean end)
            if (r == 0) return 0;
            if (r < 0) return r;
            ++frames_in;
            bytes_in += 4 + len;
            if (len == 0) return DNSPIR_E_FRAME;  // "empty request"
            switch (req[0]) {
                case CMD_INFO: {
                    ++infos;
                    const std::vector<uint8_t> o = info();
                    if ((r = write_frame(fd, o.data(), o.size()))) return r;
                    break;
                }
                case CMD_QUERY: {
                    if (!srv) {
                        ++unavailable;
                        if ((r = write_frame(fd, &RESP_UNAVAILABLE, 1))) return r;
                        break;
                    }
                    ++queries;
                    uint64_t ticket = 0;
                    if ((r = dnspir_server_submit(srv, db, client, req.data() + 1, len - 1, 0, &ticket))) return r;
                    if ((r = dnspir_server_wait(srv, ticket, reply.data(), nullptr))) return r;
                    if ((r = write_frame(fd, reply.data(), reply.size()))) return r;
                    break;
                }
                case CMD_DUMP: {
                    ++dumps;
                    if (!has_dump) {
                        ++unavailable;
                        if ((r = write_frame(fd, &RESP_UNAVAILABLE, 1))) return r;
                        break;
                    }
                    if ((r = write_frame(fd, dump.data(), dump.size()))) return r;
                    break;
                }
                default:
                    return DNSPIR_E_FRAME;  // "unknown command"
            }
        }
    }

    void accept_loop() {
        for (;;) {
            const int fd = ::accept(lfd, nullptr, nullptr);
            const int err = errno;
            std::unique_lock<std::mutex> lk(mu);
            if (stopping) {
                if (fd >= 0) ::close(fd);
                return;
            }
            if (fd < 0) {
                if (err == EMFILE || err == ENFILE) {
                    // out of descriptors: retry once connections have closed
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::milliseconds(10));
                    continue;
                }
                if (err == EINTR || err == ECONNABORTED) continue;
                return;
            }
            const int one = 1;
            (void)setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
            // reap the threads of closed connections
            for (auto it = conns.begin(); it != conns.end();) {
                if ((*it)->done) {
                    (*it)->th.join();
                    it = conns.erase(it);
                } else {
                    ++it;
                }
            }
            conns.emplace_back(new Conn);
            Conn* c = conns.back().get();
            c->fd = fd;
            c->th = std::thread([this, c] {
                if (handle(c->fd)) ++errors;  // server.rs:166-170: only this connection ends
                {
                    std::lock_guard<std::mutex> lk2(mu);
                    ::close(c->fd);
                    c->fd = -1;
                }
                c->done = true;
            });
        }
    }

    void stop() {
        std::list<std::unique_ptr<Conn>> cs;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (lfd < 0) return;
            stopping = true;
            ::shutdown(lfd, SHUT_RDWR);
            for (auto& c : conns)
                if (c->fd >= 0) ::shutdown(c->fd, SHUT_RDWR);
        }
        if (acceptor.joinable()) acceptor.join();
        {
            std::lock_guard<std::mutex> lk(mu);
            cs.swap(conns);
        }
        for (auto& c : cs)
            if (c->th.joinable()) c->th.join();
        std::lock_guard<std::mutex> lk(mu);
        ::close(lfd);
        lfd = -1;
        stopping = false;
    }
};

extern "C" {

int dnspir_frontend_create(const dnspir_frontend_config* cfg, dnspir_frontend** out) {
    if (!cfg || !out) return DNSPIR_E_ARG;
    if (cfg->dump_len && !cfg->dump) return DNSPIR_E_ARG;
    if (cfg->dump_len > FRAME_MAX || cfg->dump_len > 0xFFFFFFFFu) return DNSPIR_E_ARG;
    std::unique_ptr<dnspir_frontend> fe(new (std::nothrow) dnspir_frontend);
    if (!fe) return DNSPIR_E_OOM;
    if (cfg->server) {
        // the shape clients derive from the advertised count must be the fleet's
        int l_p = 0;
        size_t len = 0;
        if (server_db_shape(cfg->server, cfg->db, &l_p, &len)) return DNSPIR_E_ARG;
        int lp = 0, ls = 0;
        if (dnspir_database_shape(cfg->num_entries, &lp, &ls)) return DNSPIR_E_SHAPE;
        if (lp != l_p || dnspir_num_databases(cfg->num_entries) != len) return DNSPIR_E_SHAPE;
        fe->srv = cfg->server;
        fe->db = cfg->db;
        fe->num_entries = (uint32_t)cfg->num_entries;
    }
    fe->zone = cfg->zone_name ? cfg->zone_name : "";
    if (fe->zone.size() > FRAME_MAX / 2) return DNSPIR_E_ARG;
    fe->salt = cfg->hash_salt;
    fe->has_dump = cfg->dump != nullptr;
    if (cfg->dump_len) fe->dump.assign(cfg->dump, cfg->dump + cfg->dump_len);
    *out = fe.release();
    return 0;
}

int dnspir_frontend_handle_fd(dnspir_frontend* fe, int fd) {
    if (!fe || fd < 0) return DNSPIR_E_ARG;
    return fe->handle(fd);
}

int dnspir_frontend_listen(dnspir_frontend* fe, const char* host, int port, int* bound_port) {
    if (!fe || port < 0 || port > 65535) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(fe->mu);
    if (fe->lfd >= 0) return DNSPIR_E_ARG;  // already listening
    sockaddr_in a;
    memset(&a, 0, sizeof(a));
    a.sin_family = AF_INET;
    a.sin_port = htons((uint16_t)port);
    if (inet_pton(AF_INET, host ? host : "0.0.0.0", &a.sin_addr) != 1) return DNSPIR_E_ARG;
    const int fd = ::socket(AF_INET, SOCK_STREAM | SOCK_CLOEXEC, 0);
    if (fd < 0) return DNSPIR_E_IO;
    const int one = 1;
    (void)setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
    if (::bind(fd, reinterpret_cast<sockaddr*>(&a), sizeof(a)) || ::listen(fd, 128)) {
        ::close(fd);
        return DNSPIR_E_IO;
    }
    socklen_t al = sizeof(a);
    if (bound_port && ::getsockname(fd, reinterpret_cast<sockaddr*>(&a), &al) == 0) *bound_port = ntohs(a.sin_port);
    fe->lfd = fd;
    fe->acceptor = std::thread([fe] { fe->accept_loop(); });
    return 0;
}

int dnspir_frontend_get_stats(dnspir_frontend* fe, dnspir_frontend_stats* out) {
    if (!fe || !out) return DNSPIR_E_ARG;
    out->connections = fe->connections;
    out->frames_in = fe->frames_in;
    out->frames_out = fe->frames_out;
    out->bytes_in = fe->bytes_in;
    out->bytes_out = fe->bytes_out;
    out->queries = fe->queries;
    out->infos = fe->infos;
    out->dumps = fe->dumps;
    out->unavailable = fe->unavailable;
    out->errors = fe->errors;
    return 0;
}

int dnspir_frontend_destroy(dnspir_frontend* fe) {
    if (!fe) return DNSPIR_E_ARG;
    fe->stop();
    delete fe;
    return 0;
}

}  // extern "C"
