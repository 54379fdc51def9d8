// internal.h -- shared declarations of the host side of libtfhe_mi355x.
#pragma once
#include <stdint.h>

#include <string>

#include "tfhe_mi355x.h"

namespace tfhe {

// last-error text for the flat ABI (thread-local)
void set_error(const std::string& msg);
// Aborts with a message; used by the void upstream entry points, which have no error channel
// (the reference host exits on any device error, kernel/src/include/host.h:13-18).
[[noreturn]] void fatal(const char* where, int code);

// flat key export of a secret/cloud key set created by this library
void flatten_bk(const LweBootstrappingKey* bk, int32_t* out);
void flatten_ksk(const LweKeySwitchKey* ks, int32_t* out);
bool flat_params_of(const TFheGateBootstrappingParameterSet* p, tfhe_flat_params* out);

// device key bound to a cloud key set (created on first use)
tfhe_gpu_key* device_key_for(const TFheGateBootstrappingCloudKeySet* bk);
tfhe_gpu_key* device_key_for_bkfft(const TGswSampleFFT* bkfft);
tfhe_gpu_key* device_key_for_lwebk(const LweBootstrappingKeyFFT* bkfft);
tfhe_gpu_key* device_key_for_ksk(const LweKeySwitchKey* ks);
void register_keyset(const TFheGateBootstrappingCloudKeySet* bk);
void forget_device_keys(const TFheGateBootstrappingCloudKeySet* bk);
void key_params(const tfhe_gpu_key* key, tfhe_flat_params* out);
int32_t default_precision();

}  // namespace tfhe
