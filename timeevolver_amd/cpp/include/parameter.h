#pragma once
// Named model parameters written next to the results (helper/parameter.h:9-75): the same class
// names, members and paraPush macro, so a caller's parameter list carries over unchanged.
// getData() prints the value with 10 significant digits (the text used in result file names).
#include <iomanip>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

class parameter
{
public:
    parameter(const std::string& name, bool inclInFilename) : p_name(name), includeInFilename(inclInFilename) {}
    parameter() = default;
    virtual ~parameter() {}

    virtual std::string getData() = 0;
    virtual bool isDouble() = 0;
    virtual bool isInt() = 0;
    virtual bool isBool() = 0;

    std::string getName() { return p_name; }
    bool getPrintFilename() { return includeInFilename; }

private:
    std::string p_name;
    bool includeInFilename = false;
};

template <typename T>
class typedParameter : public parameter
{
public:
    typedParameter(const std::string& name, bool inclInFilename, const T& data)
        : parameter(name, inclInFilename), para_data(data) {}

    std::string getData() override
    {
        std::ostringstream s;
        s << std::setprecision(10) << para_data;
        return s.str();
    }
    bool isDouble() override { return std::is_same<T, double>::value; }
    bool isInt() override { return std::is_same<T, int>::value; }
    bool isBool() override { return std::is_same<T, bool>::value; }
    T getValue() { return para_data; }

private:
    T para_data;
};

typedef std::vector<std::shared_ptr<parameter>> parameter_list;

// paraPush("name", include_in_file_name, value): a shared typedParameter of the value's type.
#define paraPush(nm, bfnm, ...) std::make_shared<typedParameter<decltype(__VA_ARGS__)>>(nm, bfnm, __VA_ARGS__)
