// api_internal.hpp -- shared error plumbing of the libaem C ABI.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/aem.h"
#include "plan.hpp"

namespace aem {

void set_error(const std::string& m);

// Plan behind a device simulation (defined with the device state).
const Plan* sim_plan(const aem_sim* s);

// Run f, mapping exceptions to status codes and the thread-local message.
template <class F>
aem_status guard(F&& f) {
    try {
        f();
        return AEM_OK;
    } catch (const Error& e) {
        set_error(e.what());
        return (aem_status)e.status;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return AEM_ERR_STATE;
    } catch (const std::exception& e) {
        set_error(e.what());
        return AEM_ERR_STATE;
    }
}

}  // namespace aem
