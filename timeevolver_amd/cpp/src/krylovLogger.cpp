#include "krylovLogger.h"

#include <cstdio>
#include <iomanip>
#include <sstream>

#include "runtime.h"

// The severity threshold is process-global, as the Boost.Log core filter of the reference
// (core/krylovLogger.cpp:7-20).
krylovLogger::krylovLogger(loggingLevel level) { set_loggingLevel(level); }

void krylovLogger::set_loggingLevel(loggingLevel level) { te_set_log_level((int)level); }

void krylovLogger::log_message(loggingLevel level, const std::string& msg) {
    static const char* names[] = {"debug", "info", "warning", "error", "fatal"};
    std::fprintf(stderr, "[timeevolver %s] %s\n", names[(int)level], msg.c_str());
}

std::string krylovLogger::stringStreamWrapperd(double val) {
    std::stringstream s;
    s << std::scientific << std::setprecision(6) << val;
    return s.str();
}
