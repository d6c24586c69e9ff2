function [dist, idx, time, dTime, sTime] = gpuknnBitonic(query, pdata, k, numStreams)
% GPUKNNBITONIC  k nearest neighbours on the GPU, bitonic entry point.
%   [dist idx time dTime sTime] = gpuknnBitonic(query, pdata, k, numStreams)
%   query: d x q, one query per column; pdata: d x N, one corpus vector per column.
%   dist, idx: k x q squared Euclidean distances and 1-based column indices of pdata, each
%   column sorted by distance (ties: lower index first). time, dTime, sTime: seconds of the
%   call, its distance stage and its search stage. numStreams is accepted for compatibility.
%   Same calling convention as toolbox/gpuknnBitonic.m of the CUDA toolbox.
if nargin < 4
  numStreams = 1;
end
if isa(pdata, 'single')
  [dist, idx, time, dTime, sTime] = mexknnsBitonic(single(query), pdata, k, numStreams);
else
  [dist, idx, time, dTime, sTime] = mexknnsDBitonic(double(query), double(pdata), k, numStreams);
end
[dist, idx] = knns_finish(dist, idx, k);
end

function [dist, idx] = knns_finish(dist, idx, k)
% 0-based indices -> 1-based; stable sort of every column by distance; first k rows
idx = idx + 1;
[dist, order] = sort(dist, 1);
for c = 1:size(idx, 2)
  idx(:, c) = idx(order(:, c), c);
end
dist = dist(1:k, :);
idx = idx(1:k, :);
end
