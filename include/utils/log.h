/* utils/log.h -- stderr logging of the drop-in C++ API (the reference's
 * include/utils/log.h:12-26 conventions: LOG_FATAL prints and exits). */
#pragma once

#include <cstdio>
#include <cstdlib>

#define LBP_LOG(level_str, fmtstr, ...) \
    std::fprintf(stderr, "[%s] (%s:%d) " fmtstr "\n", level_str, __FILE__, __LINE__, ##__VA_ARGS__)

#define LOG_FATAL(fmtstr, ...)                     \
    do {                                           \
        LBP_LOG("FATAL", fmtstr, ##__VA_ARGS__);   \
        std::exit(EXIT_FAILURE);                   \
    } while (0)
#define LOG_ERROR(fmtstr, ...) LBP_LOG("ERROR", fmtstr, ##__VA_ARGS__)
#define LOG_WARN(fmtstr, ...) LBP_LOG("WARN", fmtstr, ##__VA_ARGS__)
#define LOG_INFO(fmtstr, ...) LBP_LOG("INFO", fmtstr, ##__VA_ARGS__)
#define LOG_DEBUG(fmtstr, ...) LBP_LOG("DEBUG", fmtstr, ##__VA_ARGS__)
