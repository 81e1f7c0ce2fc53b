// Exception-to-status translation shared by the extern "C" entry points.
#pragma once

#include <stdexcept>
#include <string>

#include "common.h"

namespace dcvc {

// Thrown for failures that map to a specific status code.
struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

template <class T>
T* checked(void* p) {
    if (!p) throw std::invalid_argument("null handle");
    return static_cast<T*>(p);
}

// Runs fn, converting exceptions into status codes (and the message into
// dcvc_last_error()).  std::invalid_argument / logic_error -> bad argument /
// call order; std::runtime_error -> malformed stream unless a StatusError.
template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return DCVC_OK;
    } catch (const StatusError& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return DCVC_ERR_ARG;
    } catch (const std::logic_error& e) {
        set_last_error(e.what());
        return DCVC_ERR_STATE;
    } catch (const std::runtime_error& e) {
        set_last_error(e.what());
        return DCVC_ERR_STREAM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return DCVC_ERR_HIP;
    } catch (...) {
        // Nothing may unwind through an extern "C" frame.
        set_last_error("unknown exception");
        return DCVC_ERR_HIP;
    }
}

}  // namespace dcvc
