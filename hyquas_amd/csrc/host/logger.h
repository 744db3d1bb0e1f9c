// logger.h -- the reference's Logger (src/logger.h:9-48): lines collected
// during compile and run, printed as "Logger: <line>" by Logger::print().
// Off by default, like the reference's SHOW_SUMMARY build option.
#pragma once

class Logger {
public:
    static void add(const char* format, ...);
    static void print();
    static void enable(bool on);
};
