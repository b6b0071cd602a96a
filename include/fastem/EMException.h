// EMException.h -- failure of the EM iteration (EM2/EMException.h:15-19).
#pragma once
#include <stdexcept>

class EMException : public std::runtime_error {
public:
    explicit EMException(const char *what) : std::runtime_error(what) {}
};
