// Grammar identity: the content hash `gagp run --save-all` writes as
// population.json's grammar_hash.
#pragma once
#include <string>

#include "json.hpp"

namespace gagp_amd {

// SHA-256 (FIPS 180-4) of `bytes`, as 64 lowercase hex digits.
std::string sha256_hex(const std::string& bytes);

// Content hash of a grammar-definition-v2 document without imports: the
// SHA-256 of the canonical JSON of its resolved form. The resolved form is
// an object with format_version, catalog_version "gagp-primitives-v3",
// normalization_version "1", resolved true, entry, search_limits,
// execution_limits, inputs and locals (empty arrays when absent),
// nonterminals and templates sorted by id (alternatives and holes sorted by
// id, each resource without its "operation" key), and source_content: the
// source document itself, with its arrays sorted the same way and no
// imports. Reference: Resolver::visit/finish and canonical_json,
// cpp/src/evolution/grammar/definition.cpp:396-511, hashed at
// cpp/src/evolution/grammar/compiled.cpp:143-145.
std::string grammar_content_hash(const json::Value& grammar_document);

}  // namespace gagp_amd
