#pragma once
// Logger with a process-global severity filter (core/krylovLogger.h:12-28). Messages go to
// stderr through the device library's logger, so the filter also applies to its messages.
#include <string>

class krylovLogger
{
public:
    enum loggingLevel { DEBUG, INFO, WARNING, ERROR, FATAL };

    krylovLogger() = default;
    krylovLogger(loggingLevel level);
    void set_loggingLevel(loggingLevel level);
    void log_message(loggingLevel level, const std::string& msg);
    static std::string stringStreamWrapperd(double val);
};
