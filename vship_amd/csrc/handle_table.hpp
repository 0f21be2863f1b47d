// handle_table.hpp - the id -> object tables behind the C ABI's handlers.
//
// Host-only (no HIP), so that the CPU tests can stress it under ThreadSanitizer.
// The reference keeps raw pointers in a vector and hands them out after
// releasing its lock (src/util/concurrency.hpp:175-200, src/VshipLib.cpp:242-258),
// so a Free on one thread can delete a handler that a Compute on another
// thread is using. Here every lookup returns a shared_ptr that the caller holds
// for the whole call: Free only drops the table's reference, and the object is
// destroyed when the last call using it returns. A freed id is marked empty and
// reused by a later insert; looking it up in between fails (BadHandler).
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace vship {

template <typename T>
class HandleTable {
public:
    // Stores obj (may be null: an init that failed still gets an id whose
    // detailed error can be read) in the first free slot; returns its id.
    unsigned insert(std::shared_ptr<T> obj, std::string last_error = std::string())
    {
        std::lock_guard<std::mutex> g(mu_);
        unsigned id = 0;
        while (id < slots_.size() && slots_[id].live) id++;
        if (id == slots_.size()) slots_.emplace_back();
        Slot& s = slots_[id];
        s.obj = std::move(obj);
        s.live = true;
        s.last_error = std::move(last_error);
        return id;
    }

    // The object of a live id (copied out under the lock), false otherwise.
    bool lookup(unsigned id, std::shared_ptr<T>* out) const
    {
        std::lock_guard<std::mutex> g(mu_);
        if (id >= slots_.size() || !slots_[id].live) return false;
        if (out) *out = slots_[id].obj;
        return true;
    }

    // Marks a live id free and hands its reference to the caller.
    bool remove(unsigned id, std::shared_ptr<T>* out)
    {
        std::lock_guard<std::mutex> g(mu_);
        if (id >= slots_.size() || !slots_[id].live) return false;
        if (out) *out = std::move(slots_[id].obj);
        slots_[id].obj.reset();
        slots_[id].live = false;
        slots_[id].last_error.clear();
        return true;
    }

    bool set_error(unsigned id, const std::string& msg)
    {
        std::lock_guard<std::mutex> g(mu_);
        if (id >= slots_.size() || !slots_[id].live) return false;
        slots_[id].last_error = msg;
        return true;
    }

    bool get_error(unsigned id, std::string* out) const
    {
        std::lock_guard<std::mutex> g(mu_);
        if (id >= slots_.size() || !slots_[id].live) return false;
        *out = slots_[id].last_error;
        return true;
    }

private:
    struct Slot {
        std::shared_ptr<T> obj;
        bool live = false;
        std::string last_error;
    };
    mutable std::mutex mu_;
    std::vector<Slot> slots_;
};

}  // namespace vship
