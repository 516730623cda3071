// S-expressions (the text form of relalg plans and VOILA programs) and the
// column catalog both front ends resolve names against.
#pragma once

#include <cctype>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"

namespace vh {

struct Sx {
	std::string atom;                 // empty for a list
	std::vector<std::shared_ptr<Sx>> list;
	bool is_list() const { return atom.empty(); }
};
typedef std::shared_ptr<Sx> SxP;

// Recursive-descent reader: atoms are runs of non-space, non-paren
// characters; ';' starts a comment to the end of the line.
struct Parser {
	const char* s;
	std::string err;
	void ws()
	{
		while (*s && (isspace((unsigned char)*s) || *s == ';')) {
			if (*s == ';')
				while (*s && *s != '\n')
					s++;
			else
				s++;
		}
	}
	SxP parse()
	{
		ws();
		if (!*s) {
			err = "unexpected end of text";
			return nullptr;
		}
		if (*s == '(') {
			s++;
			auto n = std::make_shared<Sx>();
			for (;;) {
				ws();
				if (!*s) {
					err = "missing ')'";
					return nullptr;
				}
				if (*s == ')') {
					s++;
					break;
				}
				auto c = parse();
				if (!c)
					return nullptr;
				n->list.push_back(c);
			}
			return n;
		}
		if (*s == ')') {
			err = "unexpected ')'";
			return nullptr;
		}
		auto n = std::make_shared<Sx>();
		while (*s && !isspace((unsigned char)*s) && *s != '(' && *s != ')')
			n->atom.push_back(*s++);
		return n;
	}
};

// A table of the catalog: runtime::Database's relations (Database.hpp:28-200)
// with each column's type, device data and [min, max] (Import.cpp:175-186).
struct CatalogTable {
	int64_t rows = -1;
	std::map<std::string, vh_column> cols;
};

} // namespace vh

struct vh_db {
	std::map<std::string, vh::CatalogTable> tables;
};
