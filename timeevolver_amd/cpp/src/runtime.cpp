#include "runtime.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>

namespace tecpp {

te_ctx* ctx() {
    static te_ctx* c = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* d = std::getenv("TE_DEVICE");
        const int dev = d ? std::atoi(d) : 0;
        const int rc = te_ctx_create(dev, &c);
        if (rc != TE_OK) {
            std::fprintf(stderr, "[timeevolver fatal] no usable MI355X device %d (error %d); this build has no CPU "
                                 "path\n", dev, rc);
            std::exit(1);
        }
    });
    return c;
}

void fatal(const std::string& msg) {
    std::fprintf(stderr, "[timeevolver fatal] %s\n", msg.c_str());
    std::exit(1);
}

void check(int rc, const char* what) {
    if (rc != TE_OK) fatal(std::string(what) + ": " + te_last_error(ctx()));
}

}  // namespace tecpp
