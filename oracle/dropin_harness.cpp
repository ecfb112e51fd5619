// Test infrastructure only: drives the reference-side drop-in care::mi355x::correct
// (rabbitqcplus_amd/dropin/correct_mi355x.cpp, compiled against the reference headers and
// linked to libcare_hip.so) over the harness's array-backed CpuReadStorage, so that its
// SerializedObjectStorage can be compared with the one care::cpu::correct_cpu returns.
#include "ref_harness.hpp"
#include "../rabbitqcplus_amd/dropin/correct_mi355x.hpp"

#include <correctedsequence.hpp>

#include <cstring>
#include <exception>
#include <string>

extern "C" {

// The drop-in's records in the order of its SerializedObjectStorage (read-id order)
// into the context's record buffer (ref_records_bytes / ref_records_data read them);
// returns the record count, or -1 with the exception's message in *error (256 bytes).
long long dropin_correct(ref_ctx* c, int device, char* error){
    try{
        care::SerializedObjectStorage storage = care::mi355x::correct(c->po, c->rs, device);
        c->records.clear();
        c->recordOffsets.clear();
        const std::size_t num = storage.getNumElements();
        for(std::size_t i = 0; i < num; i++){
            const std::uint8_t* ptr = storage.getPointer(i);
            std::uint32_t flags;
            std::memcpy(&flags, ptr + 4, 4);
            const std::size_t bytes = 8 + (flags & ((1u << 29) - 1));
            c->recordOffsets.push_back(c->records.size());
            c->records.insert(c->records.end(), ptr, ptr + bytes);
        }
        c->recordOffsets.push_back(c->records.size());
        return (long long)num;
    }catch(const std::exception& e){
        if(error){ std::strncpy(error, e.what(), 255); error[255] = 0; }
        return -1;
    }
}

}  // extern "C"
